// Host construction of the point-location grid of a tetrahedral field.
//
// The reference locates a point with an any-hit query of a BVH over the
// cells (sampleField, DeviceCopyableObjects.h:482-527, built in
// UnstructuredField::finalize, UnstructuredField.cpp:212-273).  A BVH query is
// a chain of ~log4(#cells) dependent node reads; here a point maps to one
// cell of a uniform grid by arithmetic, and the grid cell holds, contiguously,
// a copy of every tetrahedron that can contain a point of the cell (64 bytes
// each: four float4 {position, value}).  A lookup is one read of the cell's
// range and then independent 64-byte reads, tested in order of how much of
// the cell each tetrahedron covers, so the first test usually succeeds.
//
// Conservative by construction: the overlap test of a tetrahedron with a
// cell box uses the box grown by 1e-4 cell widths (separating-axis test in
// double precision), so every tetrahedron whose float containment test can
// pass for a point that the float cell mapping sends to this cell is listed.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace vrt {
namespace tetgrid {

struct Rec {
  float v[4][4];  // 4 vertices: x, y, z, value (v[0][3] NaN: v[1][3] is the cell value)
};

struct Grid {
  float lo[3];
  float inv_cell[3];  // cells per unit length, per axis
  int dims[3];
  std::vector<uint32_t> start;  // dims product + 1
  std::vector<Rec> recs;        // entries, cell-major
  std::vector<uint32_t> rec_tet;  // per entry: index of its tetrahedron in the input
  double mean_list = 0, expected_tests = 0;
  uint32_t max_list = 0;
};

template <class F>
inline void parallel_for(size_t n, int threads, F &&fn) {
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  threads = int(std::min<size_t>(size_t(threads), std::max<size_t>(n, 1)));
  std::atomic<size_t> next{0};
  const size_t chunk = 256;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const size_t b = next.fetch_add(chunk);
        if (b >= n) break;
        const size_t e = std::min(n, b + chunk);
        for (size_t i = b; i < e; ++i) fn(i);
      }
    });
  for (auto &th : pool) th.join();
}

// separating-axis test of a tetrahedron (double vertices) against the box
// [lo, hi]; true when no separating axis exists (they may overlap)
inline bool tet_box_overlap(const double p[4][3], const double lo[3], const double hi[3]) {
  const double c[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
  const double h[3] = {0.5 * (hi[0] - lo[0]), 0.5 * (hi[1] - lo[1]), 0.5 * (hi[2] - lo[2])};
  double q[4][3];
  for (int i = 0; i < 4; ++i)
    for (int a = 0; a < 3; ++a) q[i][a] = p[i][a] - c[a];
  auto separated = [&](const double n[3]) {
    const double r = h[0] * std::fabs(n[0]) + h[1] * std::fabs(n[1]) + h[2] * std::fabs(n[2]);
    double mn = 1e300, mx = -1e300;
    for (int i = 0; i < 4; ++i) {
      const double d = q[i][0] * n[0] + q[i][1] * n[1] + q[i][2] * n[2];
      mn = std::min(mn, d);
      mx = std::max(mx, d);
    }
    return mn > r || mx < -r;
  };
  for (int a = 0; a < 3; ++a) {
    const double n[3] = {double(a == 0), double(a == 1), double(a == 2)};
    if (separated(n)) return false;
  }
  static const int face[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  for (int f = 0; f < 4; ++f) {
    double e1[3], e2[3], n[3];
    for (int a = 0; a < 3; ++a) {
      e1[a] = q[face[f][1]][a] - q[face[f][0]][a];
      e2[a] = q[face[f][2]][a] - q[face[f][0]][a];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    if (separated(n)) return false;
  }
  static const int edge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (int e = 0; e < 6; ++e) {
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = q[edge[e][1]][a] - q[edge[e][0]][a];
    for (int a = 0; a < 3; ++a) {
      // d x unit axis a
      double n[3] = {0, 0, 0};
      if (a == 0) { n[1] = d[2]; n[2] = -d[1]; }
      if (a == 1) { n[0] = -d[2]; n[2] = d[0]; }
      if (a == 2) { n[0] = d[1]; n[1] = -d[0]; }
      if (n[0] == 0 && n[1] == 0 && n[2] == 0) continue;
      if (separated(n)) return false;
    }
  }
  return true;
}

// barycentric containment in double (for the coverage estimate only)
inline bool tet_contains(const double p[4][3], const double x[3]) {
  double m[3][3], r[3];
  for (int a = 0; a < 3; ++a) {
    for (int i = 0; i < 3; ++i) m[a][i] = p[i + 1][a] - p[0][a];
    r[a] = x[a] - p[0][a];
  }
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  if (det == 0.0) return false;
  double b[3];
  for (int i = 0; i < 3; ++i) {
    double mi[3][3];
    for (int a = 0; a < 3; ++a)
      for (int j = 0; j < 3; ++j) mi[a][j] = (j == i) ? r[a] : m[a][j];
    b[i] = (mi[0][0] * (mi[1][1] * mi[2][2] - mi[1][2] * mi[2][1]) -
            mi[0][1] * (mi[1][0] * mi[2][2] - mi[1][2] * mi[2][0]) +
            mi[0][2] * (mi[1][0] * mi[2][1] - mi[1][1] * mi[2][0])) /
           det;
  }
  return b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[0] + b[1] + b[2] <= 1;
}

// Build the grid over [lo, hi] for `tets` (records in cell-id order of the
// field).  cells_per_tet sets the resolution: about cells_per_tet * #tets
// cells, shaped like the bounds.  Returns false if the lists exceed 2^32
// entries.
inline bool build(const std::vector<Rec> &tets, const float lo[3], const float hi[3],
                  double cells_per_tet, int threads, Grid &g) {
  const size_t nt = tets.size();
  double ext[3], vol = 1.0;
  for (int a = 0; a < 3; ++a) {
    ext[a] = std::max(double(hi[a]) - double(lo[a]), 1e-30);
    vol *= ext[a];
  }
  const double target = std::max(1.0, cells_per_tet * double(std::max<size_t>(nt, 1)));
  const double edge = std::cbrt(vol / target);
  for (int a = 0; a < 3; ++a) {
    g.dims[a] = int(std::min(2048.0, std::max(1.0, std::round(ext[a] / edge))));
    g.lo[a] = lo[a];
    g.inv_cell[a] = float(double(g.dims[a]) / ext[a]);
  }
  const size_t nc = size_t(g.dims[0]) * g.dims[1] * g.dims[2];
  // the cell widths as the float mapping sees them
  double cw[3];
  for (int a = 0; a < 3; ++a) cw[a] = 1.0 / double(g.inv_cell[a]);
  auto cell_range = [&](const Rec &r, int c0[3], int c1[3]) {
    for (int a = 0; a < 3; ++a) {
      double mn = r.v[0][a], mx = r.v[0][a];
      for (int i = 1; i < 4; ++i) {
        mn = std::min(mn, double(r.v[i][a]));
        mx = std::max(mx, double(r.v[i][a]));
      }
      const double q0 = (mn - double(g.lo[a])) / cw[a] - 1e-4;
      const double q1 = (mx - double(g.lo[a])) / cw[a] + 1e-4;
      c0[a] = int(std::max(0.0, std::floor(q0)));
      c1[a] = int(std::min(double(g.dims[a] - 1), std::floor(q1)));
    }
  };
  auto overlaps = [&](const Rec &r, int x, int y, int z) {
    double p[4][3], blo[3], bhi[3];
    for (int i = 0; i < 4; ++i)
      for (int a = 0; a < 3; ++a) p[i][a] = r.v[i][a];
    const int c[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
      blo[a] = double(g.lo[a]) + (double(c[a]) - 1e-4) * cw[a];
      bhi[a] = double(g.lo[a]) + (double(c[a]) + 1.0 + 1e-4) * cw[a];
    }
    return tet_box_overlap(p, blo, bhi);
  };
  // pass 1: counts
  std::vector<std::atomic<uint32_t>> count(nc);
  for (auto &c : count) c.store(0, std::memory_order_relaxed);
  parallel_for(nt, threads, [&](size_t t) {
    int c0[3], c1[3];
    cell_range(tets[t], c0, c1);
    for (int z = c0[2]; z <= c1[2]; ++z)
      for (int y = c0[1]; y <= c1[1]; ++y)
        for (int x = c0[0]; x <= c1[0]; ++x)
          if (overlaps(tets[t], x, y, z))
            count[(size_t(z) * g.dims[1] + y) * g.dims[0] + x].fetch_add(1, std::memory_order_relaxed);
  });
  g.start.assign(nc + 1, 0);
  uint64_t total = 0;
  uint32_t mx = 0;
  for (size_t c = 0; c < nc; ++c) {
    g.start[c] = uint32_t(total);
    const uint32_t k = count[c].load(std::memory_order_relaxed);
    total += k;
    mx = std::max(mx, k);
    if (total >= (uint64_t(1) << 32)) return false;
  }
  g.start[nc] = uint32_t(total);
  g.max_list = mx;
  g.mean_list = nc ? double(total) / double(nc) : 0.0;
  // pass 2: tetrahedron ids per cell
  std::vector<uint32_t> ids(total);
  std::vector<std::atomic<uint32_t>> fill(nc);
  for (auto &c : fill) c.store(0, std::memory_order_relaxed);
  parallel_for(nt, threads, [&](size_t t) {
    int c0[3], c1[3];
    cell_range(tets[t], c0, c1);
    for (int z = c0[2]; z <= c1[2]; ++z)
      for (int y = c0[1]; y <= c1[1]; ++y)
        for (int x = c0[0]; x <= c1[0]; ++x)
          if (overlaps(tets[t], x, y, z)) {
            const size_t c = (size_t(z) * g.dims[1] + y) * g.dims[0] + x;
            ids[g.start[c] + fill[c].fetch_add(1, std::memory_order_relaxed)] = uint32_t(t);
          }
  });
  // order each list: most covered first (4^3 stratified points per cell,
  // each credited to its first containing tetrahedron by id), then by id
  g.recs.resize(total);
  g.rec_tet.resize(total);
  std::vector<double> tests(nc, 0.0);
  parallel_for(nc, threads, [&](size_t c) {
    const uint32_t b = g.start[c], e = g.start[c + 1];
    if (b == e) return;
    uint32_t *L = ids.data() + b;
    const uint32_t k = e - b;
    std::sort(L, L + k);
    std::vector<int> hits(k, 0);
    const int x = int(c % size_t(g.dims[0])), y = int((c / size_t(g.dims[0])) % size_t(g.dims[1])),
              z = int(c / (size_t(g.dims[0]) * g.dims[1]));
    const int S = 4;
    for (int sz = 0; sz < S; ++sz)
      for (int sy = 0; sy < S; ++sy)
        for (int sx = 0; sx < S; ++sx) {
          const double pt[3] = {double(g.lo[0]) + (x + (sx + 0.5) / S) * cw[0],
                                double(g.lo[1]) + (y + (sy + 0.5) / S) * cw[1],
                                double(g.lo[2]) + (z + (sz + 0.5) / S) * cw[2]};
          for (uint32_t i = 0; i < k; ++i) {
            double p[4][3];
            for (int v = 0; v < 4; ++v)
              for (int a = 0; a < 3; ++a) p[v][a] = tets[L[i]].v[v][a];
            if (tet_contains(p, pt)) {
              ++hits[i];
              break;
            }
          }
        }
    std::vector<uint32_t> order(k);
    for (uint32_t i = 0; i < k; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t bb) { return hits[a] > hits[bb]; });
    double et = 0.0;
    int covered = 0;
    for (uint32_t i = 0; i < k; ++i) {
      g.recs[b + i] = tets[L[order[i]]];
      g.rec_tet[b + i] = L[order[i]];
      et += double(i + 1) * hits[order[i]];
      covered += hits[order[i]];
    }
    // a point in no listed tetrahedron tests the whole list
    et += double(k) * double(S * S * S - covered);
    tests[c] = et / double(S * S * S);
  });
  double et = 0.0;
  for (size_t c = 0; c < nc; ++c) et += tests[c];
  g.expected_tests = nc ? et / double(nc) : 0.0;
  return true;
}

}  // namespace tetgrid
}  // namespace vrt
