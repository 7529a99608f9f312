// Seeded synthetic scene generators for the benchmark workloads and tests
// (host C++, its own small library: it produces inputs, it renders nothing).
//
// PRNG: SplitMix64, uniform doubles as (u >> 11) * 2^-53 -- portable across
// standard libraries, unlike std::uniform_*_distribution.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>
#include <unordered_map>
#include <thread>
#include <vector>

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

template <class F>
void parallel_slices(int nz, int threads, F &&fn) {
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  threads = std::min(threads, nz);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      for (int z = t; z < nz; z += threads) fn(z);
    });
  for (auto &th : pool) th.join();
}

// Improved gradient noise (Perlin 2002) over a seeded permutation.
struct Noise {
  int p[512];
  explicit Noise(uint64_t seed) {
    SplitMix64 r(seed);
    int perm[256];
    for (int i = 0; i < 256; ++i) perm[i] = i;
    for (int i = 255; i > 0; --i) {  // Fisher-Yates
      int j = int(r.next() % uint64_t(i + 1));
      std::swap(perm[i], perm[j]);
    }
    for (int i = 0; i < 512; ++i) p[i] = perm[i & 255];
  }
  static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int h, double x, double y, double z) {
    h &= 15;
    double u = h < 8 ? x : y;
    double v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
  }
  double operator()(double x, double y, double z) const {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int X = int(int64_t(fx) & 255), Y = int(int64_t(fy) & 255), Z = int(int64_t(fz) & 255);
    x -= fx;
    y -= fy;
    z -= fz;
    double u = fade(x), v = fade(y), w = fade(z);
    int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                     lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(p[AB + 1], x, y - 1, z - 1),
                          grad(p[BB + 1], x - 1, y - 1, z - 1))));
  }
};


// ---------------------------------------------------------------------------
// NanoVDB float-grid writer (file format version 32.7, the version vendored by
// the reference under external/nanovdb): one grid, one root, 32^3 upper
// internal nodes (4096^3 voxels each), 16^3 lower internal nodes (128^3),
// 8^3 leaves.  Nodes are stored breadth-first (grid | tree | root | root
// tiles | upper nodes | lower nodes | leaves); every child reference is a
// signed byte offset from the parent's data.  Record sizes and field offsets
// are those of the format (GridData 672 B, TreeData 64 B, RootData 64 B,
// root tile 32 B, upper node 270400 B, lower node 33856 B, leaf 2144 B).
// Inactive voxels of a leaf hold the background; regions without nodes read
// the background; no tiles are written.  The checksum is written as disabled
// (all bits set), as the format allows.
// ---------------------------------------------------------------------------
struct NvLeaf {
  int o[3];           // origin (multiple of 8)
  float v[512];       // (i&7)<<6 | (j&7)<<3 | (k&7)
  uint64_t mask[8];   // active voxels
};

struct Stats {
  float mn = INFINITY, mx = -INFINITY;
  double sum = 0, sum2 = 0;
  uint64_t n = 0;
  int bmin[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, bmax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  void add(float v, int i, int j, int k) {
    mn = std::min(mn, v);
    mx = std::max(mx, v);
    sum += v;
    sum2 += double(v) * v;
    ++n;
    const int c[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
      bmin[a] = std::min(bmin[a], c[a]);
      bmax[a] = std::max(bmax[a], c[a]);
    }
  }
  void add(const Stats &o) {
    if (!o.n) return;
    mn = std::min(mn, o.mn);
    mx = std::max(mx, o.mx);
    sum += o.sum;
    sum2 += o.sum2;
    n += o.n;
    for (int a = 0; a < 3; ++a) {
      bmin[a] = std::min(bmin[a], o.bmin[a]);
      bmax[a] = std::max(bmax[a], o.bmax[a]);
    }
  }
  float avg() const { return n ? float(sum / double(n)) : 0.f; }
  float dev() const {
    if (!n) return 0.f;
    const double m = sum / double(n);
    return float(std::sqrt(std::max(0.0, sum2 / double(n) - m * m)));
  }
};

template <class T> void put(uint8_t *b, size_t off, T v) { std::memcpy(b + off, &v, sizeof(T)); }

inline void put_bbox(uint8_t *b, size_t off, const Stats &s) {
  for (int a = 0; a < 3; ++a) put<int32_t>(b, off + 4 * a, s.n ? s.bmin[a] : 0);
  for (int a = 0; a < 3; ++a) put<int32_t>(b, off + 12 + 4 * a, s.n ? s.bmax[a] : 0);
}

inline void put_stats(uint8_t *b, size_t off, const Stats &s, float background) {
  put<float>(b, off, s.n ? s.mn : background);
  put<float>(b, off + 4, s.n ? s.mx : background);
  put<float>(b, off + 8, s.avg());
  put<float>(b, off + 12, s.dev());
}

constexpr size_t kGridBytes = 672, kTreeBytes = 64, kRootBytes = 64, kRootTileBytes = 32;
constexpr size_t kUpperBytes = 270400, kLowerBytes = 33856, kLeafBytes = 2144;

// Map of a uniform voxel size `s` and translation `t` (index -> world:
// world = s * index + t), float and double halves.
void put_map(uint8_t *b, size_t off, double s, const double t[3]) {
  for (int i = 0; i < 9; ++i) put<float>(b, off + 4 * i, (i % 4 == 0) ? float(s) : 0.f);
  for (int i = 0; i < 9; ++i) put<float>(b, off + 36 + 4 * i, (i % 4 == 0) ? 1.f / float(s) : 0.f);
  for (int i = 0; i < 3; ++i) put<float>(b, off + 72 + 4 * i, float(t[i]));
  put<float>(b, off + 84, 1.f);
  for (int i = 0; i < 9; ++i) put<double>(b, off + 88 + 8 * i, (i % 4 == 0) ? s : 0.0);
  for (int i = 0; i < 9; ++i) put<double>(b, off + 160 + 8 * i, (i % 4 == 0) ? 1.0 / s : 0.0);
  for (int i = 0; i < 3; ++i) put<double>(b, off + 232 + 8 * i, t[i]);
  put<double>(b, off + 256, 1.0);
}

std::vector<uint8_t> write_nanovdb(std::vector<NvLeaf> &leaves, float background, double voxel,
                                   const double trans[3], const char *name) {
  auto upper_key = [](const int o[3]) {
    // root key of a 4096^3 region: x, y, z >> 12 in 21-bit fields
    return (uint64_t(uint32_t(o[2]) >> 12)) | (uint64_t(uint32_t(o[1]) >> 12) << 21) |
           (uint64_t(uint32_t(o[0]) >> 12) << 42);
  };
  auto lower_in_upper = [](const int o[3]) {
    return uint32_t(((o[0] & 4095) >> 7) << 10 | ((o[1] & 4095) >> 7) << 5 | ((o[2] & 4095) >> 7));
  };
  auto leaf_in_lower = [](const int o[3]) {
    return uint32_t(((o[0] & 127) >> 3) << 8 | ((o[1] & 127) >> 3) << 4 | ((o[2] & 127) >> 3));
  };
  // breadth-first order: uppers by root key, lowers by (upper, slot), leaves
  // by (lower, slot)
  std::sort(leaves.begin(), leaves.end(), [&](const NvLeaf &a, const NvLeaf &b) {
    const uint64_t ka = upper_key(a.o), kb = upper_key(b.o);
    if (ka != kb) return ka < kb;
    const uint32_t la = lower_in_upper(a.o), lb = lower_in_upper(b.o);
    if (la != lb) return la < lb;
    return leaf_in_lower(a.o) < leaf_in_lower(b.o);
  });
  struct Lower { int o[3]; size_t first, count; Stats st; };
  struct Upper { int o[3]; size_t first, count; Stats st; };
  std::vector<Lower> lowers;
  std::vector<Upper> uppers;
  std::vector<Stats> leaf_stats(leaves.size());
  for (size_t i = 0; i < leaves.size(); ++i) {
    const NvLeaf &L = leaves[i];
    for (int n = 0; n < 512; ++n)
      if ((L.mask[n >> 6] >> (n & 63)) & 1ull)
        leaf_stats[i].add(L.v[n], L.o[0] + (n >> 6), L.o[1] + ((n >> 3) & 7), L.o[2] + (n & 7));
    int lo[3] = {L.o[0] & ~127, L.o[1] & ~127, L.o[2] & ~127};
    if (lowers.empty() || std::memcmp(lowers.back().o, lo, sizeof lo) != 0) {
      lowers.push_back(Lower{{lo[0], lo[1], lo[2]}, i, 0, Stats()});
      int up[3] = {lo[0] & ~4095, lo[1] & ~4095, lo[2] & ~4095};
      if (uppers.empty() || std::memcmp(uppers.back().o, up, sizeof up) != 0)
        uppers.push_back(Upper{{up[0], up[1], up[2]}, lowers.size() - 1, 0, Stats()});
      uppers.back().count++;
    }
    lowers.back().count++;
    lowers.back().st.add(leaf_stats[i]);
  }
  for (auto &u : uppers)
    for (size_t l = u.first; l < u.first + u.count; ++l) u.st.add(lowers[l].st);
  Stats root;
  for (auto &u : uppers) root.add(u.st);

  const size_t off_tree = kGridBytes, off_root = off_tree + kTreeBytes;
  const size_t off_upper = off_root + kRootBytes + uppers.size() * kRootTileBytes;
  const size_t off_lower = off_upper + uppers.size() * kUpperBytes;
  const size_t off_leaf = off_lower + lowers.size() * kLowerBytes;
  const size_t total = off_leaf + leaves.size() * kLeafBytes;
  std::vector<uint8_t> buf(total, 0);
  uint8_t *b = buf.data();

  // GridData
  put<uint64_t>(b, 0, 0x314244566f6e614eull);  // "NanoVDB1"
  put<uint64_t>(b, 8, ~0ull);                  // checksum disabled
  put<uint32_t>(b, 16, (32u << 21) | (7u << 10) | 0u);
  // flags: HasBBox | HasMinMax | HasAverage | HasStdDeviation | IsBreadthFirst
  put<uint32_t>(b, 20, root.n ? (2u | 4u | 8u | 16u | 32u) : (4u | 8u | 16u | 32u));
  put<uint32_t>(b, 24, 0u);  // grid index
  put<uint32_t>(b, 28, 1u);  // grid count
  put<uint64_t>(b, 32, uint64_t(total));
  std::strncpy(reinterpret_cast<char *>(b + 40), name ? name : "", 255);
  put_map(b, 296, voxel, trans);
  // world bbox = map([min, max + 1]) of the active index bbox
  for (int a = 0; a < 3; ++a) {
    put<double>(b, 560 + 8 * a, root.n ? voxel * root.bmin[a] + trans[a] : INFINITY);
    put<double>(b, 584 + 8 * a, root.n ? voxel * (root.bmax[a] + 1.0) + trans[a] : -INFINITY);
  }
  for (int a = 0; a < 3; ++a) put<double>(b, 608 + 8 * a, voxel);
  put<uint32_t>(b, 632, 2u);  // GridClass::FogVolume
  put<uint32_t>(b, 636, 1u);  // GridType::Float
  put<int64_t>(b, 640, int64_t(total));  // blind metadata offset: none
  put<uint32_t>(b, 648, 0u);
  put<uint64_t>(b, 664, 0x314244566f6e614eull);

  // TreeData: offsets from the tree to the first leaf, lower, upper, root
  put<int64_t>(b, off_tree + 0, int64_t(leaves.empty() ? 0 : off_leaf - off_tree));
  put<int64_t>(b, off_tree + 8, int64_t(lowers.empty() ? 0 : off_lower - off_tree));
  put<int64_t>(b, off_tree + 16, int64_t(uppers.empty() ? 0 : off_upper - off_tree));
  put<int64_t>(b, off_tree + 24, int64_t(off_root - off_tree));
  put<uint32_t>(b, off_tree + 32, uint32_t(leaves.size()));
  put<uint32_t>(b, off_tree + 36, uint32_t(lowers.size()));
  put<uint32_t>(b, off_tree + 40, uint32_t(uppers.size()));
  put<uint64_t>(b, off_tree + 56, root.n);

  // RootData + tiles (all children)
  put_bbox(b, off_root, root);
  put<uint32_t>(b, off_root + 24, uint32_t(uppers.size()));
  put<float>(b, off_root + 28, background);
  put_stats(b, off_root + 32, root, background);
  for (size_t u = 0; u < uppers.size(); ++u) {
    const size_t t = off_root + kRootBytes + u * kRootTileBytes;
    put<uint64_t>(b, t, upper_key(uppers[u].o));
    put<int64_t>(b, t + 8, int64_t(off_upper + u * kUpperBytes - off_root));
    put<uint32_t>(b, t + 16, 0u);
    put<float>(b, t + 20, background);
  }
  // internal nodes: bbox, flags, value mask (no active tiles), child mask,
  // stats, table of child offsets / background tiles
  auto internal = [&](size_t node, const Stats &st, int log2dim, size_t mask_bytes, size_t table_off) {
    put_bbox(b, node, st);
    put<uint64_t>(b, node + 24, st.n ? 2u : 0u);
    put_stats(b, node + 32 + 2 * mask_bytes, st, background);
    const size_t n = size_t(1) << (3 * log2dim);
    for (size_t i = 0; i < n; ++i) put<float>(b, node + table_off + 8 * i, background);
  };
  for (size_t u = 0; u < uppers.size(); ++u) {
    const size_t node = off_upper + u * kUpperBytes;
    internal(node, uppers[u].st, 5, 4096, 8256);
    for (size_t l = uppers[u].first; l < uppers[u].first + uppers[u].count; ++l) {
      const uint32_t n = lower_in_upper(lowers[l].o);
      b[node + 32 + 4096 + (n >> 3)] |= uint8_t(1u << (n & 7));  // child mask
      put<int64_t>(b, node + 8256 + 8 * size_t(n), int64_t(off_lower + l * kLowerBytes - node));
    }
  }
  for (size_t l = 0; l < lowers.size(); ++l) {
    const size_t node = off_lower + l * kLowerBytes;
    internal(node, lowers[l].st, 4, 512, 1088);
    for (size_t i = lowers[l].first; i < lowers[l].first + lowers[l].count; ++i) {
      const uint32_t n = leaf_in_lower(leaves[i].o);
      b[node + 32 + 512 + (n >> 3)] |= uint8_t(1u << (n & 7));
      put<int64_t>(b, node + 1088 + 8 * size_t(n), int64_t(off_leaf + i * kLeafBytes - node));
    }
  }
  // leaves: active bbox (min + extent), flags, value mask, stats, values
  for (size_t i = 0; i < leaves.size(); ++i) {
    const size_t node = off_leaf + i * kLeafBytes;
    const Stats &st = leaf_stats[i];
    for (int a = 0; a < 3; ++a) put<int32_t>(b, node + 4 * a, st.n ? st.bmin[a] : leaves[i].o[a]);
    for (int a = 0; a < 3; ++a) b[node + 12 + a] = uint8_t(st.n ? st.bmax[a] - st.bmin[a] : 0);
    b[node + 15] = st.n ? 2u : 0u;
    std::memcpy(b + node + 16, leaves[i].mask, 64);
    put_stats(b, node + 80, st, background);
    std::memcpy(b + node + 96, leaves[i].v, 2048);
  }
  return buf;
}

}  // namespace

extern "C" {

__attribute__((visibility("default"))) void vrtgen_splitmix_uniform(uint64_t seed, int n,
                                                                     double *out) {
  SplitMix64 r(seed);
  for (int i = 0; i < n; ++i) out[i] = r.uniform();
}

// Sum of `count` isotropic Gaussian blobs sampled at the voxel positions of a
// field spanning [-0.5, 0.5]^3 (voxel i at -0.5 + i/(N-1)), clamped to [0,1].
// Centres ~ U[0,1]^3 (unit cube, shifted onto the bounds), sigma ~
// U[0.03,0.12], amplitude ~ U[0.2,1], drawn in that order per blob.
__attribute__((visibility("default"))) void vrtgen_blobs(const int dims[3], int count,
                                                          uint64_t seed, float *out,
                                                          int threads) {
  SplitMix64 r(seed);
  std::vector<double> cx(count), cy(count), cz(count), s2(count), amp(count);
  for (int i = 0; i < count; ++i) {
    cx[i] = r.uniform() - 0.5;
    cy[i] = r.uniform() - 0.5;
    cz[i] = r.uniform() - 0.5;
    double sigma = 0.03 + 0.09 * r.uniform();
    s2[i] = 1.0 / (2.0 * sigma * sigma);
    amp[i] = 0.2 + 0.8 * r.uniform();
  }
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  parallel_slices(nz, threads, [&](int z) {
    double pz = -0.5 + (nz > 1 ? double(z) / (nz - 1) : 0.5);
    for (int y = 0; y < ny; ++y) {
      double py = -0.5 + (ny > 1 ? double(y) / (ny - 1) : 0.5);
      for (int x = 0; x < nx; ++x) {
        double px = -0.5 + (nx > 1 ? double(x) / (nx - 1) : 0.5);
        double v = 0;
        for (int i = 0; i < count; ++i) {
          double dx = px - cx[i], dy = py - cy[i], dz = pz - cz[i];
          v += amp[i] * std::exp(-(dx * dx + dy * dy + dz * dz) * s2[i]);
        }
        out[(size_t(z) * ny + y) * nx + x] = float(std::min(std::max(v, 0.0), 1.0));
      }
    }
  });
}

// fBm of improved gradient noise, evaluated at unit-cube voxel positions
// q = i/(N-1) scaled by `base_freq`, `octaves` octaves with lacunarity and
// gain, remapped to [0,1] by its own min/max.
__attribute__((visibility("default"))) void vrtgen_fbm(const int dims[3], int octaves,
                                                        double base_freq, double lacunarity,
                                                        double gain, uint64_t seed, float *out,
                                                        int threads) {
  Noise noise(seed);
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  std::vector<double> zmin(nz, 1e300), zmax(nz, -1e300);
  parallel_slices(nz, threads, [&](int z) {
    double qz = nz > 1 ? double(z) / (nz - 1) : 0.5;
    for (int y = 0; y < ny; ++y) {
      double qy = ny > 1 ? double(y) / (ny - 1) : 0.5;
      for (int x = 0; x < nx; ++x) {
        double qx = nx > 1 ? double(x) / (nx - 1) : 0.5;
        double f = base_freq, a = 1.0, v = 0.0;
        for (int o = 0; o < octaves; ++o) {
          v += a * noise(qx * f, qy * f, qz * f);
          f *= lacunarity;
          a *= gain;
        }
        out[(size_t(z) * ny + y) * nx + x] = float(v);
        zmin[z] = std::min(zmin[z], v);
        zmax[z] = std::max(zmax[z], v);
      }
    }
  });
  double lo = *std::min_element(zmin.begin(), zmin.end());
  double hi = *std::max_element(zmax.begin(), zmax.end());
  double inv = hi > lo ? 1.0 / (hi - lo) : 0.0;
  parallel_slices(nz, threads, [&](int z) {
    float *row = out + size_t(z) * ny * nx;
    for (size_t i = 0; i < size_t(ny) * nx; ++i) {
      double v = (double(row[i]) - lo) * inv;
      row[i] = float(std::min(std::max(v, 0.0), 1.0));
    }
  });
}

// A float NanoVDB grid of the dense x-fastest array `values` (dims[0] x
// dims[1] x dims[2], index (i, j, k) = voxel (i, j, k)): voxels that differ
// from `background` are active; map world = voxel * index + trans.
// Returns a malloc'd buffer (free with vrtgen_free) and its size.
__attribute__((visibility("default"))) int vrtgen_nanovdb_from_dense(
    const int dims[3], const float *values, float background, double voxel,
    const double trans[3], uint8_t **out, uint64_t *out_bytes) {
  std::map<std::array<int, 3>, size_t> index;
  std::vector<NvLeaf> leaves;
  for (int k = 0; k < dims[2]; ++k)
    for (int j = 0; j < dims[1]; ++j)
      for (int i = 0; i < dims[0]; ++i) {
        const float v = values[(size_t(k) * dims[1] + j) * dims[0] + i];
        if (v == background) continue;
        std::array<int, 3> o = {i & ~7, j & ~7, k & ~7};
        auto it = index.find(o);
        if (it == index.end()) {
          it = index.emplace(o, leaves.size()).first;
          leaves.emplace_back();
          NvLeaf &L = leaves.back();
          std::memcpy(L.o, o.data(), sizeof L.o);
          std::fill(L.v, L.v + 512, background);
          std::memset(L.mask, 0, sizeof L.mask);
        }
        NvLeaf &L = leaves[it->second];
        const int n = ((i & 7) << 6) | ((j & 7) << 3) | (k & 7);
        L.v[n] = v;
        L.mask[n >> 6] |= 1ull << (n & 63);
      }
  std::vector<uint8_t> buf = write_nanovdb(leaves, background, voxel, trans, "density");
  *out = static_cast<uint8_t *>(std::malloc(buf.size()));
  if (!*out) return -1;
  std::memcpy(*out, buf.data(), buf.size());
  *out_bytes = buf.size();
  return 0;
}

// Sparse fBm cloud as a NanoVDB float grid over index space [0, n)^3 with
// an identity map.  At q = (index + 1/2) / n: envelope e = 1 - |q - c| / 0.45
// (c the cube centre), g = clamp(1/2 + f/2, 0, 1) with f the fBm of
// `octaves` octaves of gradient noise at q * base_freq (lacunarity, gain),
// cloud c = e * g.  Voxels with c > tau are active with density
// (c - tau) / (1 - tau); the rest read the background 0.  tau is the
// (1 - active_fraction) quantile of c over the 32^3 sub-lattice of voxel
// centres (q = (32 m + 16) / n), so about `active_fraction` of the voxels
// are active.  Leaves whose nearest point to c has e <= tau cannot hold
// active voxels (g <= 1) and are not evaluated.
__attribute__((visibility("default"))) int vrtgen_nanovdb_cloud(
    int n, int octaves, double base_freq, double lacunarity, double gain, uint64_t seed,
    double active_fraction, int threads, uint8_t **out, uint64_t *out_bytes, double *tau_out) {
  if (n < 8 || (n & 7)) return -1;
  Noise noise(seed);
  auto cloud = [&](double qx, double qy, double qz) {
    const double dx = qx - 0.5, dy = qy - 0.5, dz = qz - 0.5;
    const double e = 1.0 - std::sqrt(dx * dx + dy * dy + dz * dz) / 0.45;
    if (e <= 0.0) return 0.0;
    double f = 0.0, fr = base_freq, a = 1.0;
    for (int o = 0; o < octaves; ++o) {
      f += a * noise(qx * fr, qy * fr, qz * fr);
      fr *= lacunarity;
      a *= gain;
    }
    const double g = std::min(std::max(0.5 + 0.5 * f, 0.0), 1.0);
    return e * g;
  };
  // threshold from the sub-lattice
  const int m = 32;
  std::vector<double> sample(size_t(m) * m * m);
  parallel_slices(m, threads, [&](int z) {
    for (int y = 0; y < m; ++y)
      for (int x = 0; x < m; ++x) {
        auto q = [&](int c) { return (double(c) * n / m + double(n) / (2 * m) + 0.5) / n; };
        sample[(size_t(z) * m + y) * m + x] = cloud(q(x), q(y), q(z));
      }
  });
  std::vector<double> sorted = sample;
  std::sort(sorted.begin(), sorted.end());
  const size_t qi = std::min(sorted.size() - 1,
                             size_t(std::max(0.0, (1.0 - active_fraction)) * double(sorted.size())));
  const double tau = std::max(sorted[qi], 1e-6);
  if (tau_out) *tau_out = tau;
  const int L = n / 8;
  std::vector<std::vector<NvLeaf>> per_slab(L);
  parallel_slices(L, threads, [&](int lz) {
    std::vector<NvLeaf> &mine = per_slab[lz];
    NvLeaf leaf;
    for (int ly = 0; ly < L; ++ly)
      for (int lx = 0; lx < L; ++lx) {
        // nearest point of the leaf's voxel-centre box to the centre
        double r2 = 0.0;
        const int lc[3] = {lx, ly, lz};
        for (int a = 0; a < 3; ++a) {
          const double lo = (lc[a] * 8 + 0.5) / n, hi = (lc[a] * 8 + 7.5) / n;
          const double d = 0.5 < lo ? lo - 0.5 : (0.5 > hi ? 0.5 - hi : 0.0);
          r2 += d * d;
        }
        if (1.0 - std::sqrt(r2) / 0.45 <= tau) continue;
        bool any = false;
        std::memset(leaf.mask, 0, sizeof leaf.mask);
        for (int i = 0; i < 8; ++i)
          for (int j = 0; j < 8; ++j)
            for (int k = 0; k < 8; ++k) {
              const int gi = lx * 8 + i, gj = ly * 8 + j, gk = lz * 8 + k;
              const double c = cloud((gi + 0.5) / n, (gj + 0.5) / n, (gk + 0.5) / n);
              const int idx = (i << 6) | (j << 3) | k;
              if (c > tau) {
                leaf.v[idx] = float((c - tau) / (1.0 - tau));
                leaf.mask[idx >> 6] |= 1ull << (idx & 63);
                any = true;
              } else {
                leaf.v[idx] = 0.f;
              }
            }
        if (!any) continue;
        leaf.o[0] = lx * 8;
        leaf.o[1] = ly * 8;
        leaf.o[2] = lz * 8;
        mine.push_back(leaf);
      }
  });
  std::vector<NvLeaf> leaves;
  size_t count = 0;
  for (auto &v : per_slab) count += v.size();
  leaves.reserve(count);
  for (auto &v : per_slab) {
    leaves.insert(leaves.end(), v.begin(), v.end());
    std::vector<NvLeaf>().swap(v);
  }
  const double zero[3] = {0, 0, 0};
  std::vector<uint8_t> buf = write_nanovdb(leaves, 0.f, 1.0, zero, "density");
  *out = static_cast<uint8_t *>(std::malloc(buf.size()));
  if (!*out) return -1;
  std::memcpy(*out, buf.data(), buf.size());
  *out_bytes = buf.size();
  return 0;
}

__attribute__((visibility("default"))) void vrtgen_free(void *p) { std::free(p); }

// Icosphere of 20 * 4^level triangles (unit icosahedron, each edge split at
// its midpoint `level` times, vertices projected to the sphere), displaced
// radially: r = radius * (1 + amp * f) with f the fBm (4 octaves, lacunarity
// 2, gain 0.5) of gradient noise at the unit-sphere position * freq + 7.
// out_vertices: (10 * 4^level + 2) float3; out_indices: 20 * 4^level uint3.
__attribute__((visibility("default"))) int vrtgen_icosphere(int level, double radius, double amp,
                                                            double freq, uint64_t seed,
                                                            float *out_vertices,
                                                            uint32_t *out_indices) {
  if (level < 0 || level > 10) return -1;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<double> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                           0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
  for (size_t i = 0; i < v.size(); i += 3) {
    const double n = std::sqrt(v[i] * v[i] + v[i + 1] * v[i + 1] + v[i + 2] * v[i + 2]);
    v[i] /= n;
    v[i + 1] /= n;
    v[i + 2] /= n;
  }
  std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1,  0, 1, 7,   0, 7, 10, 0, 10, 11, 1, 5, 9,  5, 11, 4,
                             11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2,   3, 2, 6,  3, 6, 8,
                             3, 8, 9, 4, 9, 5,  2, 4, 11, 6, 2, 10, 8, 6, 7,  9, 8, 1};
  for (int l = 0; l < level; ++l) {
    std::unordered_map<uint64_t, uint32_t> mid;
    mid.reserve(f.size());
    auto midpoint = [&](uint32_t a, uint32_t b) {
      const uint64_t key = a < b ? (uint64_t(a) << 32 | b) : (uint64_t(b) << 32 | a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      double m[3];
      for (int k = 0; k < 3; ++k) m[k] = 0.5 * (v[3 * size_t(a) + k] + v[3 * size_t(b) + k]);
      const double n = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      const uint32_t id = uint32_t(v.size() / 3);
      for (int k = 0; k < 3; ++k) v.push_back(m[k] / n);
      mid.emplace(key, id);
      return id;
    };
    std::vector<uint32_t> g;
    g.reserve(4 * f.size());
    for (size_t i = 0; i < f.size(); i += 3) {
      const uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
      const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      const uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      g.insert(g.end(), tri, tri + 12);
    }
    f.swap(g);
  }
  Noise noise(seed);
  const size_t nv = v.size() / 3;
  for (size_t i = 0; i < nv; ++i) {
    const double *p = &v[3 * i];
    double fbm = 0.0, fr = freq, a = 1.0;
    for (int o = 0; o < 4; ++o) {
      fbm += a * noise(p[0] * fr + 7.0, p[1] * fr + 7.0, p[2] * fr + 7.0);
      fr *= 2.0;
      a *= 0.5;
    }
    const double r = radius * (1.0 + amp * fbm);
    for (int k = 0; k < 3; ++k) out_vertices[3 * i + k] = float(p[k] * r);
  }
  std::memcpy(out_indices, f.data(), f.size() * sizeof(uint32_t));
  return 0;
}


// Tetrahedral mesh of a jittered lattice: (n+1)^3 vertices over [-0.5, 0.5]^3,
// vertex (i, j, k) at -0.5 + (i + d)/n per axis with d ~ U[-jitter, jitter]
// for interior lattice lines and d = 0 on the boundary planes (so the mesh
// fills the cube), three draws per vertex in x-fastest vertex order.  Each
// of the n^3 cubes is split into the 6 tetrahedra around its main diagonal
// (c0, c0+e_p0, c0+e_p0+e_p1, c0+(1,1,1) for the 6 axis permutations p),
// which matches across shared faces.  Vertex values: `blob_count` Gaussian
// blobs as vrtgen_blobs (seed + 1), at the jittered positions.  Outputs:
// positions float3, values float, index uint32 x 4 per tetrahedron (cube
// order x-fastest, permutations in lexicographic order).  Returns -1 if a
// tetrahedron's volume falls below 1% of the undistorted one (jitter too
// large), else 0.
__attribute__((visibility("default"))) int vrtgen_tet_lattice(int n, double jitter,
                                                               int blob_count, uint64_t seed,
                                                               float *positions, float *values,
                                                               uint32_t *index, int threads) {
  if (n < 1 || jitter < 0.0 || jitter >= 0.5) return -1;
  const int m = n + 1;
  const size_t nv = size_t(m) * m * m;
  SplitMix64 r(seed);
  for (size_t v = 0; v < nv; ++v) {
    const int c[3] = {int(v % m), int((v / m) % m), int(v / (size_t(m) * m))};
    for (int a = 0; a < 3; ++a) {
      const double d = (2.0 * r.uniform() - 1.0) * jitter;
      const double q = (c[a] == 0 || c[a] == n) ? double(c[a]) : double(c[a]) + d;
      positions[3 * v + a] = float(-0.5 + q / n);
    }
  }
  SplitMix64 rb(seed + 1);
  std::vector<double> cx(blob_count), cy(blob_count), cz(blob_count), s2(blob_count),
      amp(blob_count);
  for (int i = 0; i < blob_count; ++i) {
    cx[i] = rb.uniform() - 0.5;
    cy[i] = rb.uniform() - 0.5;
    cz[i] = rb.uniform() - 0.5;
    const double sigma = 0.03 + 0.09 * rb.uniform();
    s2[i] = 1.0 / (2.0 * sigma * sigma);
    amp[i] = 0.2 + 0.8 * rb.uniform();
  }
  parallel_slices(m, threads, [&](int z) {
    for (size_t v = size_t(z) * m * m; v < size_t(z + 1) * m * m; ++v) {
      const double px = positions[3 * v], py = positions[3 * v + 1], pz = positions[3 * v + 2];
      double s = 0.0;
      for (int i = 0; i < blob_count; ++i) {
        const double dx = px - cx[i], dy = py - cy[i], dz = pz - cz[i];
        s += amp[i] * std::exp(-(dx * dx + dy * dy + dz * dz) * s2[i]);
      }
      values[v] = float(std::min(std::max(s, 0.0), 1.0));
    }
  });
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  std::atomic<int> bad{0};
  const double vmin = 0.01 / (6.0 * double(n) * n * n);
  parallel_slices(n, threads, [&](int k) {
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const size_t cube = (size_t(k) * n + j) * n + i;
        for (int t = 0; t < 6; ++t) {
          int c[3] = {i, j, k};
          uint32_t *out = index + 4 * (6 * cube + t);
          out[0] = uint32_t((size_t(c[2]) * m + c[1]) * m + c[0]);
          for (int s = 0; s < 2; ++s) {
            ++c[perm[t][s]];
            out[s + 1] = uint32_t((size_t(c[2]) * m + c[1]) * m + c[0]);
          }
          out[3] = uint32_t((size_t(k + 1) * m + (j + 1)) * m + (i + 1));
          double e[3][3];
          for (int s = 0; s < 3; ++s)
            for (int a = 0; a < 3; ++a)
              e[s][a] = double(positions[3 * out[s + 1] + a]) - double(positions[3 * out[0] + a]);
          const double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                             e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                             e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
          if (std::fabs(det) / 6.0 < vmin) bad.fetch_add(1);
        }
      }
  });
  return bad.load() ? -1 : 0;
}

}  // extern "C"
