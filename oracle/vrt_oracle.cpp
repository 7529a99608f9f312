/*
 * ORACLE -- test infrastructure, not product code.  See vrt_oracle.h.
 *
 * A CPU restatement of anari-visionaray's volume rendering path written
 * for checking the HIP kernels; it mirrors the reference CPU device's
 * execution structure (a std::thread pool of hardware_concurrency() threads
 * working on 64x64 pixel tiles, for_each.h:55-87, SyncContext.cpp:20-21) so
 * that it also serves as the timed CPU baseline.
 */
#include "vrt_oracle.h"
#include "nvdb_oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>
#include <map>
#include <array>

namespace orc {

// Russian roulette on (the reference) or off (orc_set_roulette, a test switch)
static bool g_roulette = true;

// ---------------------------------------------------------------------------
// small vector type, templated on the arithmetic precision
// ---------------------------------------------------------------------------
template <class R> struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a) : x(a), y(a), z(a) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  R &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / length(a)); }
template <class R> inline R max_element(V3<R> a) { return std::max(a.x, std::max(a.y, a.z)); }
template <class R> inline R min_element(V3<R> a) { return std::min(a.x, std::min(a.y, a.z)); }
template <class R> inline V3<R> vabs(V3<R> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class R> inline V3<R> from(const float *p) { return {R(p[0]), R(p[1]), R(p[2])}; }

template <class R> struct V4 { R x, y, z, w; };

// ---------------------------------------------------------------------------
// A1: LCG<4> with TEA seeding (common.h:55-102)
// ---------------------------------------------------------------------------
struct Random {
  uint32_t state;
  Random() : state(0) {}
  Random(uint32_t v0_, uint32_t v1_) {
    // TEA-4 seed, common.h:70-82
    uint32_t v0 = v0_, v1 = v1_, s0 = 0;
    for (int n = 0; n < 4; n++) {
      s0 += 0x9e3779b9u;
      v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
      v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    state = v0;
  }
  // common.h:85-91: state = 1664525*state + 1013904223; (state & 0xFFFFFF)/2^24
  float operator()() {
    state = 1664525u * state + 1013904223u;
    return float(state & 0x00FFFFFFu) / float(0x01000000);
  }
};

// ---------------------------------------------------------------------------
// Rays, boxes
// ---------------------------------------------------------------------------
template <class R> struct Ray {
  V3<R> ori, dir;
  R tmin, tmax;
};

template <class R> struct BoxHit { bool hit; R tnear, tfar; };

// visionaray intersect(ray, aabb) slab test (parity unpinned; textbook slab
// method: tnear = max of per-axis minima, tfar = min of per-axis maxima,
// hit = tfar >= tnear && tfar >= 0)
template <class R>
inline BoxHit<R> intersect_box(const Ray<R> &ray, V3<R> lo, V3<R> hi) {
  V3<R> inv(R(1) / ray.dir.x, R(1) / ray.dir.y, R(1) / ray.dir.z);
  V3<R> t1 = (lo - ray.ori) * inv;
  V3<R> t2 = (hi - ray.ori) * inv;
  R tn = std::max(std::min(t1.z, t2.z), std::max(std::min(t1.y, t2.y), std::min(t1.x, t2.x)));
  R tf = std::min(std::max(t1.z, t2.z), std::min(std::max(t1.y, t2.y), std::max(t1.x, t2.x)));
  BoxHit<R> h;
  h.tnear = tn;
  h.tfar = tf;
  h.hit = tf >= tn && tf >= R(0);
  return h;
}

// x^y from IEEE basic operations (fused multiply-add, ldexp, rint) and a
// table: the opacity correction powf(1 - alpha, dt/unitDistance)
// (VolumeIntegration.h:416-417) evaluated so that it rounds identically on
// every machine that follows these steps.  log2: x = 2^e z with z in
// [0.69921875, 1.3984375) (float bits offset by 0x3f330000); z's subinterval
// i (the next 6 mantissa bits) holds c_i (its midpoint, or exactly 1 for the
// subinterval holding 1.0); the table keeps 1/c_i rounded to float and
// -log2 of that float, r = fma(z, 1/c_i, -1), |r| <= 2^-7, and log2(1 + r)
// is its Taylor series to r^4.  2^(y log2 x) = ldexp(2^f, n), n = rint, f the
// rest, 2^f from its Taylor series to f^7.  Results below 2^-126 are 0.
// Error <= 1.5 (2 + |y log2 x|) ulp against std::pow, pinned in tests, as is
// the table.  This is parity mode (float); oracle mode (double) calls
// std::pow.
static const float kPowLog2[64][2] = {
{1.4222222566604614f, -0.5081469416618347f},
    {1.4065934419631958f, -0.4922053813934326f},
    {1.39130437374115f, -0.47643807530403137f},
    {1.376344084739685f, -0.46084117889404297f},
    {1.3617020845413208f, -0.4454111158847809f},
    {1.3473684787750244f, -0.43014445900917053f},
    {1.3333333730697632f, -0.4150375425815582f},
    {1.3195875883102417f, -0.4000871181488037f},
    {1.3061224222183228f, -0.38529011607170105f},
    {1.2929292917251587f, -0.37064337730407715f},
    {1.2799999713897705f, -0.3561437726020813f},
    {1.2673267126083374f, -0.34178850054740906f},
    {1.2549020051956177f, -0.3275747001171112f},
    {1.2427184581756592f, -0.31349948048591614f},
    {1.2307692766189575f, -0.2995603382587433f},
    {1.2190476655960083f, -0.285754531621933f},
    {1.2075471878051758f, -0.27207955718040466f},
    {1.1962616443634033f, -0.25853297114372253f},
    {1.185185194015503f, -0.24511250853538513f},
    {1.174311876296997f, -0.23181560635566711f},
    {1.163636326789856f, -0.21864023804664612f},
    {1.1531531810760498f, -0.20558416843414307f},
    {1.1428571939468384f, -0.1926451474428177f},
    {1.1327433586120605f, -0.17982102930545807f},
    {1.1228070259094238f, -0.16710999608039856f},
    {1.1130434274673462f, -0.15450988709926605f},
    {1.1034482717514038f, -0.14201900362968445f},
    {1.0940171480178833f, -0.12963534891605377f},
    {1.0847457647323608f, -0.11735695600509644f},
    {1.0756303071975708f, -0.10518231242895126f},
    {1.0666667222976685f, -0.09310948103666306f},
    {1.0578511953353882f, -0.08113670349121094f},
    {1.049180269241333f, -0.06926257908344269f},
    {1.0406503677368164f, -0.057485442608594894f},
    {1.0322580337524414f, -0.04580364748835564f},
    {1.0240000486373901f, -0.034215785562992096f},
    {1.0158730745315552f, -0.022720159962773323f},
    {1.0078740119934082f, -0.011315307579934597f},
    {1.0f, -0.0f},
    {0.9846153855323792f, 0.0223678108304739f},
    {0.9696969985961914f, 0.044394075870513916f},
    {0.9552238583564758f, 0.0660892203450203f},
    {0.9411764740943909f, 0.08746283501386642f},
    {0.9275362491607666f, 0.10852442681789398f},
    {0.9142857193946838f, 0.12928301095962524f},
    {0.9014084339141846f, 0.14974714815616608f},
    {0.8888888955116272f, 0.16992498934268951f},
    {0.8767123222351074f, 0.18982456624507904f},
    {0.8648648858070374f, 0.20945332944393158f},
    {0.8533333539962769f, 0.22881865501403809f},
    {0.8421052694320679f, 0.2479275017976761f},
    {0.8311688303947449f, 0.2667865455150604f},
    {0.8205128312110901f, 0.28540220856666565f},
    {0.8101266026496887f, 0.3037807047367096f},
    {0.800000011920929f, 0.32192808389663696f},
    {0.790123462677002f, 0.33984997868537903f},
    {0.7804877758026123f, 0.35755205154418945f},
    {0.7710843086242676f, 0.3750394880771637f},
    {0.761904776096344f, 0.39231738448143005f},
    {0.7529411911964417f, 0.4093908965587616f},
    {0.7441860437393188f, 0.42626476287841797f},
    {0.7356321811676025f, 0.4429435133934021f},
    {0.7272727489471436f, 0.45943158864974976f},
    {0.7191011309623718f, 0.4757334291934967f}};

inline float pow_ieee(float x, float y) {
  if (y == 0.f || x == 1.f) return 1.f;
  if (!(x > 0.f)) return y > 0.f ? 0.f : std::numeric_limits<float>::infinity();
  uint32_t bits;
  std::memcpy(&bits, &x, 4);
  if (bits < 0x00800000u) return y > 0.f ? 0.f : std::numeric_limits<float>::infinity();
  if (bits >= 0x7f800000u) return x;
  // exponents 0.5 and 2 (the usual sampling rates): one correctly rounded
  // operation each
  if (y == 0.5f) return std::sqrt(x);
  if (y == 2.f) {
    const float s = x * x;
    return s < std::numeric_limits<float>::min() ? 0.f : s;
  }
  const uint32_t tmp = bits - 0x3f330000u;
  const int e = int32_t(tmp) >> 23;
  const float *t = kPowLog2[(tmp >> 17) & 63u];
  const uint32_t zb = bits - (tmp & 0xff800000u);
  float z;
  std::memcpy(&z, &zb, 4);
  const float r = std::fma(z, t[0], -1.f);
  float q = std::fma(r, -0.36067376f, 0.48089835f);  // -1/(4 ln2), 1/(3 ln2)
  q = std::fma(r, q, -0.72134752f);                   // -1/(2 ln2)
  q = std::fma(r, q, 1.44269504f);                    // 1/ln2
  const float l2 = std::fma(r, q, float(e) + t[1]);
  const float zz = y * l2;
  if (zz < -126.f) return 0.f;
  if (zz > 128.f) return std::numeric_limits<float>::infinity();
  const float n = std::rint(zz);
  const float f = zz - n;
  float p = 1.52527338e-5f;            // ln2^7 / 7!
  p = std::fma(p, f, 1.54035304e-4f);  // ln2^6 / 6!
  p = std::fma(p, f, 1.33335581e-3f);
  p = std::fma(p, f, 9.61812911e-3f);
  p = std::fma(p, f, 5.55041087e-2f);
  p = std::fma(p, f, 0.240226507f);
  p = std::fma(p, f, 0.693147181f);
  p = std::fma(p, f, 1.f);
  return std::ldexp(p, int(n));
}

// the table, for its pin (tests)
extern "C" void orc_pow_log2_table(float *out) {
  for (int i = 0; i < 64; ++i) {
    out[2 * i] = kPowLog2[i][0];
    out[2 * i + 1] = kPowLog2[i][1];
  }
}

inline float opacity_pow(float x, float y) { return pow_ieee(x, y); }
inline double opacity_pow(double x, double y) { return std::pow(x, y); }

// renderer/common.h:14-18
template <class R> inline R epsilon_from(V3<R> P, V3<R> dir, R t) {
  const R ulpEpsilon = R(0x1.fp-18);
  R m = std::max(max_element(vabs(P)), max_element(vabs(dir)) * t);
  return m * ulpEpsilon;
}

// ---------------------------------------------------------------------------
// A12: unstructured field of tetrahedra (UnstructuredField.cpp:29-198,
// sampleField DeviceCopyableObjects.h:482-527, intersectTet Plane.h:56-96)
// ---------------------------------------------------------------------------
// intersectTet: vertices relative to P; makePlane(a, b, c) has N = cross(b-a,
// c-a), d = dot(a, N), eval(v) = dot(v, N) - d; the weight of a vertex is
// eval(0)/eval(vertex) of the plane through the other three; the tests are
// fa in [0,1], then fb, fc, fd >= 0 (the reference checks "fa > 1" again
// where "fb > 1" etc. would be expected, Plane.h:81-89)
template <class R> inline R tet_plane_ratio(V3<R> a, V3<R> b, V3<R> c, V3<R> opp) {
  const V3<R> n = cross(b - a, c - a);
  const R d = dot(a, n);
  return (dot(V3<R>(R(0)), n) - d) / (dot(opp, n) - d);
}

template <class R>
inline bool intersect_tet(R &value, V3<R> P, V3<R> a, R wa, V3<R> b, R wb, V3<R> c, R wc,
                          V3<R> d, R wd) {
  const V3<R> va = a - P, vb = b - P, vc = c - P, vd = d - P;
  const R fa = tet_plane_ratio(vb, vd, vc, va);
  if (fa < R(0) || fa > R(1)) return false;
  const R fb = tet_plane_ratio(va, vc, vd, vb);
  if (fb < R(0) || fa > R(1)) return false;
  const R fc = tet_plane_ratio(va, vd, vb, vc);
  if (fc < R(0) || fa > R(1)) return false;
  const R fd = tet_plane_ratio(va, vb, vc, vd);
  if (fd < R(0) || fa > R(1)) return false;
  value = fa * wa + fb * wb + fc * wc + fd * wd;
  return true;
}

struct TetField {
  struct Cell {
    uint32_t v[4];
    uint32_t src;  // index in the ANARI cell arrays: the primitive id
    float cell_value;
    float lo[3], hi[3];
    float vmin, vmax;
  };
  const float *pos = nullptr;
  std::vector<float> w;  // vertex values, NaN without vertex data
  std::vector<Cell> cells;
  float blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
  float cell_size = 0.f;
  int64_t skipped = 0;
  // the oracle's own point-location aid: a uniform grid over the bounds
  // listing, in ascending cell id, every cell whose box meets the grid cell
  int G = 1;
  std::vector<uint32_t> gstart, gids;

  // finalize (UnstructuredField.cpp:60-170): cells with fewer than 4
  // indices or a type/count mismatch are skipped; only tetrahedra are
  // restated (false if another kind of cell remains)
  bool open(const orc_volume *v) {
    pos = v->vertex_position;
    w.resize(size_t(v->num_vertices));
    for (int64_t i = 0; i < v->num_vertices; ++i)
      w[i] = v->vertex_data ? v->vertex_data[i] : std::numeric_limits<float>::quiet_NaN();
    for (int a = 0; a < 3; ++a) {
      blo[a] = FLT_MAX;
      bhi[a] = -FLT_MAX;
    }
    float avg = 0.f;
    for (int64_t c = 0; c < v->num_cells; ++c) {
      const uint64_t first = v->cell_index[c];
      const uint64_t last = c < v->num_cells - 1 ? uint64_t(v->cell_index[c + 1])
                                                 : uint64_t(v->num_indices);
      const uint64_t ic = last - first;
      const uint8_t t = v->cell_type[c];
      const bool tet = t == 0 || t == 10, hex = t == 1 || t == 12, wedge = t == 2 || t == 13,
                 pyr = t == 3 || t == 14;
      if (ic < 4 || (tet && ic != 4) || (pyr && ic != 5) || (wedge && ic != 6) ||
          (hex && ic != 8)) {
        ++skipped;
        continue;
      }
      if (ic != 4) return false;
      Cell cell;
      cell.src = uint32_t(c);
      cell.cell_value = v->cell_data ? v->cell_data[c] : 0.f;
      for (int a = 0; a < 3; ++a) {
        cell.lo[a] = FLT_MAX;
        cell.hi[a] = -FLT_MAX;
      }
      cell.vmin = FLT_MAX;
      cell.vmax = -FLT_MAX;
      for (int k = 0; k < 4; ++k) {
        cell.v[k] = v->index[first + k];
        for (int a = 0; a < 3; ++a) {
          cell.lo[a] = std::min(cell.lo[a], pos[3 * size_t(cell.v[k]) + a]);
          cell.hi[a] = std::max(cell.hi[a], pos[3 * size_t(cell.v[k]) + a]);
        }
        // buildGrid's value range (UnstructuredField.cpp:387-399)
        const float wk = w[cell.v[k]];
        const float val = std::isnan(wk) ? cell.cell_value : wk;
        cell.vmin = std::min(cell.vmin, val);
        cell.vmax = std::max(cell.vmax, val);
      }
      cells.push_back(cell);
      for (int a = 0; a < 3; ++a) {
        blo[a] = std::min(blo[a], cell.lo[a]);
        bhi[a] = std::max(bhi[a], cell.hi[a]);
      }
      // avgCellDiagonal += length(max - min) / m_elements.size(), the
      // size counting the cells kept so far (UnstructuredField.cpp:170)
      const float dx = cell.hi[0] - cell.lo[0], dy = cell.hi[1] - cell.lo[1],
                  dz = cell.hi[2] - cell.lo[2];
      avg += std::sqrt(dx * dx + dy * dy + dz * dz) / float(cells.size());
    }
    cell_size = avg;
    G = std::max(1, int(std::cbrt(double(cells.size()) / 4.0)));
    const size_t ng = size_t(G) * G * G;
    std::vector<std::vector<uint32_t>> lists(ng);
    for (size_t id = 0; id < cells.size(); ++id) {
      int c0[3], c1[3];
      for (int a = 0; a < 3; ++a) {
        c0[a] = grid_coord(a, double(cells[id].lo[a]));
        c1[a] = grid_coord(a, double(cells[id].hi[a]));
      }
      for (int z = c0[2]; z <= c1[2]; ++z)
        for (int y = c0[1]; y <= c1[1]; ++y)
          for (int x = c0[0]; x <= c1[0]; ++x)
            lists[(size_t(z) * G + y) * G + x].push_back(uint32_t(id));
    }
    gstart.assign(ng + 1, 0);
    for (size_t i = 0; i < ng; ++i) gstart[i + 1] = gstart[i] + uint32_t(lists[i].size());
    gids.reserve(gstart[ng]);
    for (auto &l : lists) gids.insert(gids.end(), l.begin(), l.end());
    return !cells.empty();
  }

  // monotone in x, so a box [lo, hi] containing x covers x's grid cell
  int grid_coord(int a, double x) const {
    const double q = (x - double(blo[a])) / (double(bhi[a]) - double(blo[a])) * G;
    return std::min(std::max(int(std::floor(std::max(q, -1.0))), 0), G - 1);
  }

  template <class R> bool test_cell(size_t id, V3<R> P, R &value) const {
    const Cell &c = cells[id];
    V3<R> p[4];
    R wv[4];
    for (int k = 0; k < 4; ++k) {
      p[k] = from<R>(pos + 3 * size_t(c.v[k]));
      wv[k] = R(w[c.v[k]]);
    }
    if (!intersect_tet(value, P, p[0], wv[0], p[1], wv[1], p[2], wv[2], p[3], wv[3]))
      return false;
    // no vertex data: the cell value (DeviceCopyableObjects.h:150-153)
    if (std::isnan(w[c.v[0]])) value = R(c.cell_value);
    return true;
  }

  // sampleField: the first cell (in id order) containing P, found through
  // the grid; false outside every cell
  template <class R> bool sample(V3<R> P, R &value, int64_t *prim = nullptr) const {
    for (int a = 0; a < 3; ++a)
      if (!(P[a] >= R(blo[a]) && P[a] <= R(bhi[a]))) return false;
    const int x = grid_coord(0, double(P.x)), y = grid_coord(1, double(P.y)),
              z = grid_coord(2, double(P.z));
    const size_t g = (size_t(z) * G + y) * G + x;
    for (uint32_t i = gstart[g]; i < gstart[g + 1]; ++i)
      if (test_cell(gids[i], P, value)) {
        if (prim) *prim = cells[gids[i]].src;
        return true;
      }
    return false;
  }

  template <class R> bool sample_brute(V3<R> P, R &value) const {
    for (size_t id = 0; id < cells.size(); ++id)
      if (test_cell(id, P, value)) return true;
    return false;
  }
};

// ---------------------------------------------------------------------------
// Volume: field + TF (A10, A13)
// ---------------------------------------------------------------------------
template <class R> struct Volume {
  const orc_volume *v;
  NanoVDBGrid nv;      // nanovdb field
  bool is_nvdb = false;
  TetField tets;       // unstructured field
  bool is_tets = false;
  V3<R> glo, gho;      // majorant grid bounds in voxel space
  V3<R> blo, bhi;      // StructuredRegularField::bounds, :90-93
  V3<R> S, T;          // voxelSpaceTransform = (scaling(1/size), -min), :37-40
  R cellSize;          // min(spacing), :40
  // instance transform
  R A[9];              // affineInv
  V3<R> transInv;      // -translation
  bool has_grid;

  explicit Volume(const orc_volume *vol) : v(vol) {
    V3<R> o = from<R>(vol->origin), sp = from<R>(vol->spacing);
    // bounds computed in float like the reference (float3 math)
    float bl[3], bh[3];
    for (int i = 0; i < 3; ++i) {
      bl[i] = vol->origin[i];
      bh[i] = vol->origin[i] + (float(vol->dims[i]) - 1.f) * vol->spacing[i];
    }
    blo = from<R>(bl);
    bhi = from<R>(bh);
    S = V3<R>(R(1) / (bhi.x - blo.x), R(1) / (bhi.y - blo.y), R(1) / (bhi.z - blo.z));
    T = -blo;
    cellSize = std::min(sp.x, std::min(sp.y, sp.z));
    (void)o;
    // invert the 3x3 part of the instance transform
    const float *m = vol->xfm;
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], k = m[10];
    double det = a * (e * k - f * h) - b * (d * k - f * g) + c * (d * h - e * g);
    double inv[9] = {(e * k - f * h) / det, (c * h - b * k) / det, (b * f - c * e) / det,
                     (f * g - d * k) / det, (a * k - c * g) / det, (c * d - a * f) / det,
                     (d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det};
    for (int i = 0; i < 9; ++i) A[i] = R(inv[i]);
    transInv = V3<R>(R(-m[3]), R(-m[7]), R(-m[11]));
    has_grid = vol->majorants != nullptr && vol->grid_dims[0] > 0;
    glo = V3<R>(R(0));
    gho = V3<R>(R(1));
    if (vol->field_type == 1) {
      // NanoVDBField::finalize (NanoVDBField.cpp:43-60): float world bbox,
      // identity voxel-space transform, cell size 1, grid over the bounds
      is_nvdb = nv.open(vol->nanovdb);
      float bl[3], bh[3];
      for (int i = 0; i < 3; ++i) {
        bl[i] = float(nv.wlo[i]);
        bh[i] = float(nv.whi[i]);
      }
      blo = from<R>(bl);
      bhi = from<R>(bh);
      S = V3<R>(R(1));
      T = V3<R>(R(0));
      cellSize = R(1);
      glo = blo;
      gho = bhi;
    }
    if (vol->field_type == 2) {
      // UnstructuredField::finalize (UnstructuredField.cpp:280-285): identity
      // voxel-space transform, bounds of the cells, the running cell size,
      // grid over the bounds
      is_tets = tets.open(vol);
      blo = from<R>(tets.blo);
      bhi = from<R>(tets.bhi);
      S = V3<R>(R(1));
      T = V3<R>(R(0));
      cellSize = R(tets.cell_size);
      glo = blo;
      gho = bhi;
    }
  }

  // the normal transform transpose(affineInv) (Instance.cpp:96-98) applied
  // as mat3 * vec3 = col0 n.x + col1 n.y + col2 n.z
  V3<R> mul_normal(V3<R> n) const {
    return {A[0] * n.x + A[3] * n.y + A[6] * n.z, A[1] * n.x + A[4] * n.y + A[7] * n.z,
            A[2] * n.x + A[5] * n.y + A[8] * n.z};
  }
  V3<R> mulA(V3<R> p) const {
    return {A[0] * p.x + A[1] * p.y + A[2] * p.z, A[3] * p.x + A[4] * p.y + A[5] * p.z,
            A[6] * p.x + A[7] * p.y + A[8] * p.z};
  }
  // Instance intersect: DeviceCopyableObjects.h:1772-1814
  Ray<R> to_local(const Ray<R> &r) const {
    Ray<R> l = r;
    l.ori = mulA(r.ori + transInv);
    l.dir = mulA(r.dir);
    return l;
  }
  // SpatialField::pointToVoxelSpace, DeviceCopyableObjects.h:390-395
  V3<R> point_to_voxel(V3<R> p) const { return S * (p + T); }
  V3<R> vector_to_voxel(V3<R> d) const { return S * d; }

  R voxel(int x, int y, int z) const {
    return R(v->voxels[(size_t(z) * v->dims[1] + y) * v->dims[0] + x]);
  }

  // tex3D(linear, clamp, normalised coordinates): texel centres at
  // (i+0.5)/N, i.e. texel coordinate p*N - 0.5 (GridAccel.cpp:103 samples
  // the TF at (i+.5)/N under the same convention)
  // sampleField (DeviceCopyableObjects.h:474-570): false where the field has
  // no value (a point outside every cell of an unstructured field)
  // sampleGradient (DeviceCopyableObjects.h:582-597) at voxel-space P with
  // delta = cellSize * diag(voxelSpaceTransform), normalised; no normal
  // (0) unless all six samples have a value and the gradient is not zero
  // (the reference's normalize(0) would be NaN)
  V3<R> gradient_normal(V3<R> P) const {
    const V3<R> delta = S * cellSize;
    R x0, x1, y0, y1, z0, z1;
    const bool b0 = sample(P + V3<R>(delta.x, R(0), R(0)), x1);
    const bool b1 = sample(P - V3<R>(delta.x, R(0), R(0)), x0);
    const bool b2 = sample(P + V3<R>(R(0), delta.y, R(0)), y1);
    const bool b3 = sample(P - V3<R>(R(0), delta.y, R(0)), y0);
    const bool b4 = sample(P + V3<R>(R(0), R(0), delta.z), z1);
    const bool b5 = sample(P - V3<R>(R(0), R(0), delta.z), z0);
    if (!(b0 && b1 && b2 && b3 && b4 && b5)) return V3<R>(R(0));
    const V3<R> g = V3<R>(x1, y1, z1) - V3<R>(x0, y0, z0);
    if (g.x == R(0) && g.y == R(0) && g.z == R(0)) return V3<R>(R(0));
    return normalize(g);
  }

  // `prim`: the primitive id of the sample -- the containing cell of an
  // unstructured field, 0 for the other fields (:478-481, 556-567)
  bool sample(V3<R> P, R &value, int64_t *prim = nullptr) const {
    if (prim) *prim = 0;
    if (v->field_type == 2) return is_tets && tets.sample(P, value, prim);
    value = sample_regular(P);
    return true;
  }

  R sample_regular(V3<R> P) const {
    if (is_nvdb) {
      const R p[3] = {P.x, P.y, P.z};
      return nv.sample(p, v->filter == 1 ? 0 : 1);
    }
    int n[3] = {v->dims[0], v->dims[1], v->dims[2]};
    int i0[3], i1[3];
    R w[3];
    for (int a = 0; a < 3; ++a) {
      R c = P[a] * R(n[a]) - R(0.5);
      c = std::min(std::max(c, R(-1)), R(n[a]));  // keep the int cast defined
      R fl = std::floor(c);
      w[a] = c - fl;
      int i = int(fl);
      i0[a] = std::min(std::max(i, 0), n[a] - 1);
      i1[a] = std::min(std::max(i + 1, 0), n[a] - 1);
    }
    // a + t (b - a) with one rounding for the product and sum (fused)
    auto lerp = [](R a, R b, R t) { return std::fma(t, b - a, a); };
    R c00 = lerp(voxel(i0[0], i0[1], i0[2]), voxel(i1[0], i0[1], i0[2]), w[0]);
    R c10 = lerp(voxel(i0[0], i1[1], i0[2]), voxel(i1[0], i1[1], i0[2]), w[0]);
    R c01 = lerp(voxel(i0[0], i0[1], i1[2]), voxel(i1[0], i0[1], i1[2]), w[0]);
    R c11 = lerp(voxel(i0[0], i1[1], i1[2]), voxel(i1[0], i1[1], i1[2]), w[0]);
    R c0 = lerp(c00, c10, w[1]);
    R c1 = lerp(c01, c11, w[1]);
    return lerp(c0, c1, w[2]);
  }

  // postClassify, DeviceCopyableObjects.h:615-619: normalise by valueRange,
  // then tex1D(linear, clamp) over the float4 table
  V4<R> classify(R val) const {
    R lo = R(v->value_lo), hi = R(v->value_hi);
    R u = (val - lo) / (hi - lo);
    int n = v->tf_size;
    R c = u * R(n) - R(0.5);
    c = std::min(std::max(c, R(-1)), R(n));
    R fl = std::floor(c);
    R w = c - fl;
    int i = int(fl);
    int i0 = std::min(std::max(i, 0), n - 1), i1 = std::min(std::max(i + 1, 0), n - 1);
    const float *t0 = v->tf_rgba + 4 * i0, *t1 = v->tf_rgba + 4 * i1;
    V4<R> r;
    r.x = R(t0[0]) + w * (R(t1[0]) - R(t0[0]));
    r.y = R(t0[1]) + w * (R(t1[1]) - R(t0[1]));
    r.z = R(t0[2]) + w * (R(t1[2]) - R(t0[2]));
    r.w = R(t0[3]) + w * (R(t1[3]) - R(t0[3]));
    return r;
  }

  R majorant(int cell) const {
    // GridAccel maxOpacities, or 1 without a grid (DeviceCopyableObjects.h:724)
    return has_grid ? R(v->majorants[cell]) : R(1);
  }
};

// ---------------------------------------------------------------------------
// A8: dda3 (renderer/DDA.h:14-112), GridAccel-common.h:10-24
// ---------------------------------------------------------------------------
template <class R, class F>
void dda3(Ray<R> ray, const int dims[3], V3<R> lo, V3<R> hi, F &&func) {
  const R ray_tmin = ray.tmin;
  ray.ori = ray.ori + ray.dir * ray.tmin;
  ray.tmin = R(0);
  ray.tmax -= ray_tmin;

  V3<R> rcp(R(1) / ray.dir.x, R(1) / ray.dir.y, R(1) / ray.dir.z);
  V3<R> l = (lo - ray.ori) * rcp;
  V3<R> h = (hi - ray.ori) * rcp;
  V3<R> tnear(std::min(l.x, h.x), std::min(l.y, h.y), std::min(l.z, h.z));
  V3<R> tfar(std::max(l.x, h.x), std::max(l.y, h.y), std::max(l.z, h.z));

  // projectOnGrid, GridAccel-common.h:18-24
  int cell[3];
  for (int a = 0; a < 3; ++a) {
    R v01 = (ray.ori[a] - lo[a]) / (hi[a] - lo[a]);
    R c = v01 * R(dims[a]);
    c = std::min(std::max(c, R(-1)), R(dims[a]));
    int ci = int(c);  // truncation as in vec3i(V01*dims)
    cell[a] = std::min(std::max(ci, 0), dims[a] - 1);
  }

  R dist[3], tnext[3];
  int step[3], stop[3];
  for (int a = 0; a < 3; ++a) {
    R d = ray.dir[a];
    step[a] = d > R(0) ? 1 : -1;
    stop[a] = d > R(0) ? dims[a] : -1;
    if (d == R(0)) {
      // DDA.h:32-40 sets tnear = FLT_MAX on zero components: the axis never
      // becomes the closest boundary
      dist[a] = R(0);
      tnext[a] = std::numeric_limits<R>::infinity();
      continue;
    }
    dist[a] = std::max(R(0), (tfar[a] - tnear[a]) / R(dims[a]));
    tnext[a] = d > R(0) ? tnear[a] + R(cell[a] + 1) * dist[a]
                        : tnear[a] + R(dims[a] - cell[a]) * dist[a];
  }

  R t0 = R(0);
  while (true) {
    R tc = std::min(tnext[0], std::min(tnext[1], tnext[2]));
    R t1 = std::min(tc, ray.tmax);
    int li = (cell[2] * dims[1] + cell[1]) * dims[0] + cell[0];
    if (!func(li, ray_tmin + t0, ray_tmin + t1)) return;
    // DDA.h:88-108: advance every axis whose boundary equals the closest
    bool out = false;
    for (int a = 0; a < 3 && !out; ++a) {
      if (tnext[a] == tc) {
        tnext[a] += dist[a];
        cell[a] += step[a];
        if (cell[a] == stop[a]) out = true;
      }
    }
    if (out) break;
    t0 = t1;
  }
}

// ---------------------------------------------------------------------------
// A7: intersect(Ray, Volume) -- box clip, voxel space, DDA + Woodcock
// (DeviceCopyableObjects.h:680-796); with `ratio`, ratio tracking
// ---------------------------------------------------------------------------
template <class R> struct VolHit {
  bool hit = false;
  R t = std::numeric_limits<R>::max();
  V3<R> albedo;
  R Tr = R(1);
  int64_t prim = 0;  // primitive id of the collision's sample
};

struct Counter { uint64_t lookups = 0; };

// Delta tracking (ratio = false): the first accepted collision, Tr = 0 at
// it.  Ratio tracking (ratio = true, an extension for shadow and AO rays:
// Novak et al. 2014, "Residual ratio tracking"): the same free-flight steps
// over the same majorant cells, but every tentative collision multiplies the
// transmittance estimate Tr by 1 - sigma/majorant instead of drawing an
// acceptance, and the walk runs to the end of the segment; E[Tr] is the
// transmittance, as the reference's binary Tr estimates it
// (DeviceCopyableObjects.h:722-757, Pathtrace_impl.cpp:441-444).
template <class R>
VolHit<R> delta_track(const Volume<R> &vol, Ray<R> ray, Random &rnd, Counter *cnt,
                      bool ratio = false) {
  VolHit<R> out;
  BoxHit<R> bh = intersect_box(ray, vol.blo, vol.bhi);
  if (!bh.hit) return out;

  ray.tmin = std::max(ray.tmin, bh.tnear);
  ray.tmax = std::min(ray.tmax, bh.tfar);
  ray.ori = vol.point_to_voxel(ray.ori);
  ray.dir = vol.vector_to_voxel(ray.dir);
  const R dt_scale = length(ray.dir);
  ray.dir = normalize(ray.dir);
  ray.tmin = ray.tmin * dt_scale;
  ray.tmax = ray.tmax * dt_scale;
  R unitDistance = R(vol.v->unit_distance) * dt_scale;
  R invUnitDistance = R(1) / unitDistance;

  bool hit = false;
  R hit_t = ray.tmax;
  V3<R> albedo;
  int64_t hit_prim = 0;
  R tr = R(1);

  auto woodcock = [&](int leafID, R t0, R t1) -> bool {
    const R majorant = leafID >= 0 ? vol.majorant(leafID) : R(1);
    R t = t0;
    while (true) {
      if (majorant <= R(0)) break;
      t -= std::log(R(1) - R(rnd())) / (majorant * invUnitDistance);
      if (t >= t1) break;
      V3<R> P = ray.ori + ray.dir * t;
      R v;
      int64_t prim = 0;
      if (cnt) cnt->lookups++;
      // the acceptance draw only where the field has a value (:739-753)
      if (vol.sample(P, v, &prim)) {
        V4<R> s = vol.classify(v);
        if (ratio) {
          tr = tr * (R(1) - s.w / majorant);
          continue;
        }
        R u = R(rnd());
        if (s.w >= u * majorant) {
          albedo = V3<R>(s.x, s.y, s.z);
          hit = true;
          hit_t = t;
          hit_prim = prim;
          return false;
        }
      }
    }
    return true;
  };

  if (vol.has_grid) {
    int dims[3] = {vol.v->grid_dims[0], vol.v->grid_dims[1], vol.v->grid_dims[2]};
    // gridBounds = voxel-space image of the world bounds = [0,1]^3
    // (StructuredRegularField.cpp:103-107)
    dda3(ray, dims, vol.glo, vol.gho, woodcock);
  } else {
    woodcock(-1, ray.tmin, ray.tmax);
  }

  if (hit) {
    out.hit = true;
    out.t = hit_t / dt_scale;
    out.albedo = albedo;
    out.Tr = R(0);
    out.prim = hit_prim;
  } else {
    out.Tr = tr;
  }
  return out;
}

// ---------------------------------------------------------------------------
// Camera (visionaray's thin_lens_camera is not vendored; the pinhole ray is
// pinned against the textbook look-at / perspective projection and the thin
// lens against its geometry, in the tests).
// ANARI perspective camera: position, direction, up, fovy, aspect,
// imageRegion; pixel (x, y) samples screen point ((x+.5)/W, (y+.5)/H), row 0
// at the bottom of the image.  Lens radius 0 consumes no random numbers.
// ---------------------------------------------------------------------------
template <class R> struct Camera {
  V3<R> pos, U, V, W;
  // thin lens: radius / |U| and / |V| (the lens plane is spanned by U, V)
  R lens_radius, lens_u, lens_v, focus_distance;
  R region[4];
  R shutter[2];
  explicit Camera(const orc_camera *c) {
    V3<R> d = normalize(from<R>(c->dir));
    V3<R> up = normalize(from<R>(c->up));
    R th = std::tan(R(c->fovy) * R(0.5));
    V3<R> right = normalize(cross(d, up));
    V3<R> upv = cross(right, d);
    pos = from<R>(c->pos);
    U = right * (th * R(c->aspect));
    V = upv * th;
    W = d;
    lens_radius = R(c->aperture_radius);
    lens_u = lens_radius > R(0) ? lens_radius / length(U) : R(0);
    lens_v = lens_radius > R(0) ? lens_radius / length(V) : R(0);
    focus_distance = R(c->focus_distance);
    for (int i = 0; i < 4; ++i) region[i] = R(c->region[i]);
    shutter[0] = R(c->shutter[0]);
    shutter[1] = R(c->shutter[1]);
  }
  Ray<R> primary(R x, R y, R width, R height) const {
    R sx = (x + R(0.5)) / width, sy = (y + R(0.5)) / height;
    sx = (R(1) - sx) * region[0] + sx * region[2];
    sy = (R(1) - sy) * region[1] + sy * region[3];
    V3<R> d = W + U * (R(2) * sx - R(1)) + V * (R(2) * sy - R(1));
    Ray<R> r;
    r.ori = pos;
    r.dir = normalize(d);
    r.tmin = R(0);
    r.tmax = std::numeric_limits<float>::max();
    return r;
  }
};

// ---------------------------------------------------------------------------
// Sampling helpers (parity unpinned; PBRT definitions)
// ---------------------------------------------------------------------------
template <class R> inline V3<R> uniform_sample_sphere(R u1, R u2) {
  R z = R(1) - R(2) * u1;
  R r = std::sqrt(std::max(R(0), R(1) - z * z));
  R phi = R(2 * M_PI) * u2;
  return V3<R>(r * std::cos(phi), r * std::sin(phi), z);
}
// Henyey-Greenstein phase function (an extension, no reference counterpart;
// g = 0 is the isotropic phase above): hg_4pi = 4 pi p_HG(cos theta) =
// (1 - g^2) / (1 + g^2 - 2 g cos theta)^(3/2); samples by the inverse CDF
// cos theta = (1 + g^2 - ((1 - g^2) / (1 - g + 2 g u1))^2) / (2 g) around
// the unit direction d, phi = 2 pi u2, in the orthonormal basis of Duff et
// al., "Building an Orthonormal Basis, Revisited" (2017).  g != 0.
template <class R> inline R hg_4pi(R g, R cosT) {
  const R d = R(1) + g * g - R(2) * g * cosT;
  return (R(1) - g * g) / (d * std::sqrt(d));
}
template <class R> inline V3<R> hg_sample(R g, V3<R> d, R u1, R u2) {
  const R s = (R(1) - g * g) / (R(1) - g + R(2) * g * u1);
  const R cosT = std::min(std::max((R(1) + g * g - s * s) / (R(2) * g), R(-1)), R(1));
  const R sinT = std::sqrt(std::max(R(0), R(1) - cosT * cosT));
  const R phi = R(2 * M_PI) * u2;
  const R sign = std::copysign(R(1), d.z);
  const R a = R(-1) / (sign + d.z);
  const R b = d.x * d.y * a;
  const V3<R> t1(R(1) + sign * d.x * d.x * a, sign * b, -sign * d.x);
  const V3<R> t2(b, sign + d.y * d.y * a, -d.y);
  return t1 * (sinT * std::cos(phi)) + t2 * (sinT * std::sin(phi)) + d * cosT;
}
// Shirley-Chiu concentric map of [0,1)^2 onto the unit disk
template <class R> inline void concentric_disk(R u1, R u2, R &dx, R &dy) {
  R sx = R(2) * u1 - R(1), sy = R(2) * u2 - R(1);
  dx = 0, dy = 0;
  if (!(sx == R(0) && sy == R(0))) {
    R r, theta;
    if (std::fabs(sx) > std::fabs(sy)) {
      r = sx;
      theta = R(M_PI / 4) * (sy / sx);
    } else {
      r = sy;
      theta = R(M_PI / 2) - R(M_PI / 4) * (sx / sy);
    }
    dx = r * std::cos(theta);
    dy = r * std::sin(theta);
  }
}
template <class R> inline V3<R> cosine_sample_hemisphere(R u1, R u2) {
  // the concentric disk, lifted to the hemisphere
  R dx, dy;
  concentric_disk(u1, u2, dx, dy);
  R z = std::sqrt(std::max(R(0), R(1) - dx * dx - dy * dy));
  return V3<R>(dx, dy, z);
}
// Thin lens (camera/Perspective.cpp:42-43, apertureRadius / focusDistance;
// PBRT 3rd ed. 6.2.3): the pinhole ray's point on the plane in focus,
// focus_distance along the view direction, seen from the lens point
// off = lens radius * concentric_disk(u1, u2) in the plane of U and V; the
// new direction dir * ft - off is that point minus the new origin ori + off
template <class R> inline void thin_lens(const Camera<R> &c, Ray<R> &r, R u1, R u2) {
  R dx, dy;
  concentric_disk(u1, u2, dx, dy);
  const R ft = c.focus_distance / dot(r.dir, c.W);
  const V3<R> off = c.U * (c.lens_u * dx) + c.V * (c.lens_v * dy);
  r.dir = normalize(r.dir * ft - off);
  r.ori = r.ori + off;
}
// ONB as in the matte sampler, renderer/common.h:1384-1386
template <class R> inline void make_onb(V3<R> w, V3<R> &u, V3<R> &v) {
  v = std::fabs(w.x) > std::fabs(w.y) ? normalize(V3<R>(-w.z, R(0), w.x))
                                      : normalize(V3<R>(R(0), w.z, -w.y));
  u = cross(v, w);
}
// faceforward(n, i, nref): n flipped into the hemisphere of i
template <class R> inline V3<R> faceforward(V3<R> n, V3<R> i, V3<R> nref) {
  return dot(i, nref) < R(0) ? -n : n;
}

// Majorants are stored as 8-bit multiples of step = (max TF opacity)/255,
// rounded up (a rounded-up majorant keeps delta tracking unbiased); the
// kernels keep the bytes in LDS.  Same float operations as the GPU build.
inline float majorant_step(const orc_volume *v) {
  float tf_max = 0.f;
  for (int i = 0; i < v->tf_size; ++i) tf_max = std::max(tf_max, v->tf_rgba[4 * i + 3]);
  float step = tf_max / 255.f;
  if (step * 255.f < tf_max) step = std::nextafter(step, std::numeric_limits<float>::infinity());
  return step;
}

inline float quantize_majorant(float m, float step) {
  if (!(m > 0.f)) return 0.f;
  int q = int(std::min(255.f, std::ceil(m / step)));
  while (q < 255 && float(q) * step < m) ++q;
  return float(q) * step;
}

// ---------------------------------------------------------------------------
// Scene wrapper
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// A6: triangles (visionaray basic_triangle {v1, e1, e2}, not vendored: parity
// unpinned).  Moller-Trumbore with inclusive edges; a hit needs
// tmin <= t <= tmax.  Closest-hit ties go to the lower triangle index so the
// result does not depend on the acceleration structure.
// ---------------------------------------------------------------------------
template <class R> struct Tri {
  V3<R> v1, e1, e2;
};

template <class R> struct TriHit {
  bool hit = false;
  R t = R(0), u = R(0), v = R(0);
};

template <class R> inline TriHit<R> intersect_tri(const Ray<R> &ray, const Tri<R> &tri) {
  TriHit<R> h;
  const V3<R> s1 = cross(ray.dir, tri.e2);
  const R div = dot(s1, tri.e1);
  if (div == R(0)) return h;
  const R inv = R(1) / div;
  const V3<R> d = ray.ori - tri.v1;
  const R b1 = dot(d, s1) * inv;
  if (b1 < R(0) || b1 > R(1)) return h;
  const V3<R> s2 = cross(d, tri.e1);
  const R b2 = dot(ray.dir, s2) * inv;
  if (b2 < R(0) || b1 + b2 > R(1)) return h;
  h.t = dot(tri.e2, s2) * inv;
  h.u = b1;
  h.v = b2;
  h.hit = true;
  return h;
}

// A5 (surfaces): a binary BVH over one mesh, median split of the centroids
// along the longest axis, leaves of <= 4 triangles -- a builder unrelated to
// the GPU's, so equal closest hits pin the GPU traversal.
template <class R> struct Mesh {
  std::vector<Tri<R>> tris;  // in input order (prim id = index)
  V3<R> color;
  struct Node {
    V3<R> lo, hi;
    int left = -1, right = -1, first = 0, count = 0;
  };
  std::vector<Node> nodes;
  std::vector<int> order;

  const orc_trimesh *m;
  explicit Mesh(const orc_trimesh *mesh) : m(mesh) {
    color = from<R>(m->color);
    tris.resize(size_t(m->num_triangles));
    for (int i = 0; i < m->num_triangles; ++i) {
      uint32_t idx[3] = {uint32_t(3 * i), uint32_t(3 * i + 1), uint32_t(3 * i + 2)};
      if (m->indices)
        for (int k = 0; k < 3; ++k) idx[k] = m->indices[3 * i + k];
      // basic_triangle edges formed in float from the vertex positions
      float v[3][3];
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) v[k][a] = m->vertices[3 * idx[k] + a];
      float e1[3], e2[3];
      for (int a = 0; a < 3; ++a) {
        e1[a] = v[1][a] - v[0][a];
        e2[a] = v[2][a] - v[0][a];
      }
      tris[i] = Tri<R>{from<R>(v[0]), from<R>(e1), from<R>(e2)};
    }
    order.resize(tris.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
    if (!tris.empty()) build(0, int(tris.size()));
  }
  void bounds(int first, int count, V3<R> &lo, V3<R> &hi) const {
    lo = V3<R>(std::numeric_limits<R>::max());
    hi = V3<R>(-std::numeric_limits<R>::max());
    for (int i = first; i < first + count; ++i) {
      const Tri<R> &t = tris[order[i]];
      const V3<R> p[3] = {t.v1, t.v1 + t.e1, t.v1 + t.e2};
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) {
          lo[a] = std::min(lo[a], p[k][a]);
          hi[a] = std::max(hi[a], p[k][a]);
        }
    }
  }
  int build(int first, int count) {
    const int id = int(nodes.size());
    nodes.emplace_back();
    V3<R> lo, hi;
    bounds(first, count, lo, hi);
    nodes[id].lo = lo;
    nodes[id].hi = hi;
    if (count <= 4) {
      nodes[id].first = first;
      nodes[id].count = count;
      return id;
    }
    const V3<R> ext = hi - lo;
    const int axis = ext.x >= ext.y && ext.x >= ext.z ? 0 : (ext.y >= ext.z ? 1 : 2);
    auto centre = [&](int i) {
      const Tri<R> &t = tris[i];
      return t.v1[axis] * R(3) + t.e1[axis] + t.e2[axis];
    };
    const int mid = first + count / 2;
    std::nth_element(order.begin() + first, order.begin() + mid, order.begin() + first + count,
                     [&](int a, int b) { return centre(a) < centre(b); });
    const int l = build(first, mid - first);
    const int r = build(mid, first + count - mid);
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
  }
  static bool box_hit(const Ray<R> &ray, V3<R> lo, V3<R> hi, R tmin, R tmax) {
    for (int a = 0; a < 3; ++a) {
      const R inv = R(1) / ray.dir[a];
      R t0 = (lo[a] - ray.ori[a]) * inv, t1 = (hi[a] - ray.ori[a]) * inv;
      if (t0 > t1) std::swap(t0, t1);
      if (!(t0 <= t1)) {  // NaN: the origin lies on a slab plane of a zero component
        if (ray.ori[a] < lo[a] || ray.ori[a] > hi[a]) return false;
        continue;
      }
      tmin = std::max(tmin, t0);
      tmax = std::min(tmax, t1);
    }
    // slack for rounding of the box itself
    return tmin <= tmax * R(1.0000002) + R(1e-30);
  }
  // closest hit in [tmin, tmax] (any hit when `any`); prim = triangle index
  bool intersect(const Ray<R> &ray, bool any, R &t_best, int &prim, R &u, R &v) const {
    if (nodes.empty()) return false;
    bool found = false;
    t_best = ray.tmax;
    int stack[128], sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node &n = nodes[stack[--sp]];
      if (!box_hit(ray, n.lo, n.hi, ray.tmin, t_best)) continue;
      if (n.count) {
        for (int i = n.first; i < n.first + n.count; ++i) {
          const int p = order[i];
          const TriHit<R> h = intersect_tri(ray, tris[p]);
          if (!h.hit || h.t < ray.tmin || h.t > t_best) continue;
          if (found && h.t == t_best && p > prim) continue;
          found = true;
          t_best = h.t;
          prim = p;
          u = h.u;
          v = h.v;
          if (any) return true;
        }
      } else {
        stack[sp++] = n.left;
        stack[sp++] = n.right;
      }
    }
    return found;
  }
};

// A17: quad light (DeviceCopyableObjects.h:2187-2273): two triangles
// {v1, e1, e1+e2} and {v1, e1+e2, e2}
template <class R> struct QuadLight {
  Tri<R> t1, t2;
  V3<R> intensity;  // color * intensity (area_light::intensity, parity unpinned)
  R A1 = R(0), A2 = R(0);
  explicit QuadLight(const orc_light &L) {
    const V3<R> v1 = from<R>(L.pos), e1 = from<R>(L.e1), e2 = from<R>(L.e2);
    t1 = Tri<R>{v1, e1, e1 + e2};
    t2 = Tri<R>{v1, e1 + e2, e2};
    intensity = from<R>(L.color) * R(L.intensity);
    A1 = R(0.5) * length(cross(t1.e1, t1.e2));
    A2 = R(0.5) * length(cross(t2.e1, t2.e2));
  }
  V3<R> normal() const { return normalize(cross(t1.e1, t1.e2)); }
  // intersect(Ray, Quad): the closer of the two triangles, t > 0
  TriHit<R> intersect(const Ray<R> &ray) const {
    TriHit<R> h1 = intersect_tri(ray, t1), h2 = intersect_tri(ray, t2), r;
    if (h1.hit && h1.t > R(0)) r = h1;
    if (h2.hit && h2.t > R(0) && (!r.hit || h2.t < r.t)) r = h2;
    return r;
  }
};

// HDRI environment light (HDRI.cpp:31-162, sampleCDF.h,
// DeviceCopyableObjects.h:2398-2418).  The map coordinates and lookups are
// float, with the M_PI terms of toPolar / toUV in double as written there.
struct HdriLight {
  int w = 0, h = 0;
  std::vector<float> rgb, rows, cols;
  float scale = 1.f;
  float W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, L[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  bool ok = false;

  explicit HdriLight(const orc_light &l) {
    if (l.type != ORC_LIGHT_HDRI || !l.hdri_rgb || l.hdri_w < 1 || l.hdri_h < 1) return;
    w = l.hdri_w;
    h = l.hdri_h;
    scale = l.scale;
    rgb.assign(l.hdri_rgb, l.hdri_rgb + size_t(w) * h * 3);
    // makeCDF (:31-66): per row the running sum of max(r, g, b), divided by
    // the row total unless it is 0; the totals' running sum likewise
    cols.resize(size_t(w) * h);
    rows.resize(h);
    for (int y = 0; y < h; ++y) {
      float *r = cols.data() + size_t(y) * w;
      for (int x = 0; x < w; ++x) {
        const float *c = rgb.data() + 3 * (size_t(y) * w + x);
        r[x] = std::max(std::max(c[0], c[1]), c[2]);
        if (x > 0) r[x] = r[x - 1] + r[x];
      }
      rows[y] = r[w - 1];
      const float last = r[w - 1];
      if (last != 0.f)
        for (int x = 0; x < w; ++x) r[x] /= last;
    }
    for (int y = 1; y < h; ++y) rows[y] = rows[y - 1] + rows[y];
    const float last = rows[h - 1];
    if (last != 0.f)
      for (int y = 0; y < h; ++y) rows[y] /= last;
    // the frame (:150-155): col1 = -up, col0 = col1 x direction,
    // col2 = col0 x col1, normalised; the inverse by the adjugate
    auto nrm = [](V3<float> a) { return a * (1.f / length(a)); };
    const V3<float> c1 = -nrm(from<float>(l.up));
    const V3<float> c0 = nrm(cross(c1, from<float>(l.dir)));
    const V3<float> c2 = nrm(cross(c0, c1));
    const V3<float> cs[3] = {c0, c1, c2};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) W[3 * r + c] = cs[c][r];
    const float det = W[0] * (W[4] * W[8] - W[5] * W[7]) - W[1] * (W[3] * W[8] - W[5] * W[6]) +
                      W[2] * (W[3] * W[7] - W[4] * W[6]);
    if (!(det != 0.f) || !std::isfinite(det)) return;
    const float inv = 1.f / det;
    L[0] = (W[4] * W[8] - W[5] * W[7]) * inv;
    L[1] = (W[2] * W[7] - W[1] * W[8]) * inv;
    L[2] = (W[1] * W[5] - W[2] * W[4]) * inv;
    L[3] = (W[5] * W[6] - W[3] * W[8]) * inv;
    L[4] = (W[0] * W[8] - W[2] * W[6]) * inv;
    L[5] = (W[2] * W[3] - W[0] * W[5]) * inv;
    L[6] = (W[3] * W[7] - W[4] * W[6]) * inv;
    L[7] = (W[1] * W[6] - W[0] * W[7]) * inv;
    L[8] = (W[0] * W[4] - W[1] * W[3]) * inv;
    ok = true;
  }

  // mat3 * vec3 = col0 x + col1 y + col2 z
  static V3<float> mul(const float *M, V3<float> d) {
    return {M[0] * d.x + M[1] * d.y + M[2] * d.z, M[3] * d.x + M[4] * d.y + M[5] * d.z,
            M[6] * d.x + M[7] * d.y + M[8] * d.z};
  }
  // sample_cdf (:40-54): std::upper_bound, the bin's probability (0 past
  // the end, which the reference leaves unset)
  static unsigned bin(const float *d, unsigned n, float x, float &pdf) {
    const unsigned i = unsigned(std::upper_bound(d, d + n, x) - d);
    pdf = i == 0 ? d[0] : (i < n ? d[i] - d[i - 1] : 0.f);
    return i;
  }
  // sampleCDF (:88-98)
  void cdf(float rx, float ry, unsigned &x, unsigned &y, float &pdfx, float &pdfy) const {
    y = std::min(bin(rows.data(), unsigned(h), ry, pdfy), unsigned(h - 1));
    x = bin(cols.data() + size_t(y) * w, unsigned(w), rx, pdfx);
  }
  // toPolar (:58-72), toUV (:74-85)
  static V3<float> to_polar(float u, float v) {
    const float theta = float(2.0 * M_PI * double(u) + -M_PI / 2.0);
    const float phi = float(M_PI * double(v));
    const float sp = std::sin(phi);
    return {-(std::cos(theta) * sp), std::cos(phi), std::sin(theta) * sp};
  }
  static void to_uv(V3<float> n, float &u, float &v) {
    const float a = std::atan2(n.x, n.z);
    u = float((double(a) + M_PI / 2.0) / (M_PI * 2.0) + M_PI * double(28.670f / 360.0f));
    // fmin / fmax: a NaN (|n.y| rounded past 1) clamps to .001, as on the GPU
    v = std::fmin(std::fmax(float(double(std::acos(n.y)) / M_PI), .001f), .999f);
  }
  // intensity (:2413-2417): bilinear, clamp, texel centres (parity
  // unpinned: visionaray's tex2D), times scale
  V3<float> radiance(V3<float> dir) const {
    float u, v;
    to_uv(mul(L, dir), u, v);
    auto axis = [](float c, int n, int &i0, int &i1) {
      c = std::min(std::max(c, -1.f), float(n));
      const float fl = std::floor(c);
      i0 = std::min(std::max(int(fl), 0), n - 1);
      i1 = std::min(std::max(int(fl) + 1, 0), n - 1);
      return c - fl;
    };
    int i0, i1, j0, j1;
    const float wx = axis(u * float(w) - 0.5f, w, i0, i1);
    const float wy = axis(v * float(h) - 0.5f, h, j0, j1);
    auto t = [&](int i, int j, int k) { return rgb[3 * (size_t(j) * w + i) + k]; };
    auto lp = [](float a, float b, float s) { return a + s * (b - a); };
    V3<float> r;
    for (int k = 0; k < 3; ++k)
      r[k] = lp(lp(t(i0, j0, k), t(i1, j0, k), wx), lp(t(i0, j1, k), t(i1, j1, k), wx), wy);
    return r * scale;
  }
  // asHDRI.sample: rx then ry (drawn left to right), the cell corner's
  // direction, pdf = pdfx pdfy w h / 4 pi
  void sample(Random &rnd, V3<float> &dir, float &pdf) const {
    const float rx = rnd(), ry = rnd();
    unsigned x, y;
    float pdfx, pdfy;
    cdf(rx, ry, x, y, pdfx, pdfy);
    const float invjac = float(unsigned(w) * unsigned(h)) / float(4.0 * M_PI);
    dir = mul(W, to_polar(float(x) / float(w), float(y) / float(h)));
    pdf = pdfx * pdfy * invjac;
  }
  // the light pdf of a direction that met the map (Pathtrace_impl.cpp:160-172)
  float pdf(V3<float> raydir) const {
    const V3<float> d = mul(L, raydir);
    float u, v;
    to_uv(d, u, v);
    unsigned x, y;
    float pdfx, pdfy;
    cdf(u, v, x, y, pdfx, pdfy);
    const float sinT = std::sin(std::acos(std::min(std::max(d.y, -1.f), 1.f)));
    if (sinT == 0.f) return 0.f;
    const float pi = float(M_PI);
    return (pdfx * pdfy) * float(unsigned(w) * unsigned(h)) / (2.0f * pi * pi * sinT);
  }
};

// An instance's transforms (orc_instance), row-major 3x3 products formed as
// (m0 x + m1 y) + m2 z
template <class R> struct InstXf {
  R A[9], tinv[3], N[9];
  explicit InstXf(const orc_instance &I) {
    for (int i = 0; i < 9; ++i) A[i] = R(I.affine_inv[i]);
    for (int i = 0; i < 3; ++i) tinv[i] = R(I.trans_inv[i]);
    for (int i = 0; i < 9; ++i) N[i] = R(I.normal_xfm[i]);
  }
  static V3<R> mul(const R *M, V3<R> p) {
    return {M[0] * p.x + M[1] * p.y + M[2] * p.z, M[3] * p.x + M[4] * p.y + M[5] * p.z,
            M[6] * p.x + M[7] * p.y + M[8] * p.z};
  }
  // intersect(Ray, Instance) (DeviceCopyableObjects.h:1796-1798)
  Ray<R> to_object(const Ray<R> &r) const {
    Ray<R> l = r;
    l.ori = mul(A, r.ori + V3<R>(tinv[0], tinv[1], tinv[2]));
    l.dir = mul(A, r.dir);
    return l;
  }
};

template <class R> struct Scene {
  std::vector<Volume<R>> vols;
  std::vector<Mesh<R>> meshes;
  std::vector<Mesh<R>> imeshes;     // per instance
  std::vector<InstXf<R>> ixf;       // per instance
  std::vector<QuadLight<R>> quads;  // per light (unused entries for other types)
  std::vector<HdriLight> hdris;     // per light (empty for other types)
  const orc_scene *s;
  explicit Scene(const orc_scene *sc) : s(sc) {
    for (int i = 0; i < sc->num_volumes; ++i) vols.emplace_back(&sc->volumes[i]);
    for (int i = 0; i < sc->num_meshes; ++i) meshes.emplace_back(&sc->meshes[i]);
    for (int i = 0; i < sc->num_instances; ++i) {
      imeshes.emplace_back(&sc->instances[i].mesh);
      ixf.emplace_back(sc->instances[i]);
    }
    for (int i = 0; i < sc->num_lights; ++i) quads.emplace_back(sc->lights[i]);
    for (int i = 0; i < sc->num_lights; ++i) hdris.emplace_back(sc->lights[i]);
  }
  bool empty() const {
    return s->num_volumes == 0 && s->num_meshes == 0 && s->num_instances == 0;
  }
  // surface `id`: mesh id, or instance id - meshes.size()
  const Mesh<R> &surface(int id) const {
    return id < int(meshes.size()) ? meshes[size_t(id)] : imeshes[size_t(id) - meshes.size()];
  }
  // its triangle's geometric normal in world space: normalize(e1 x e2), then
  // normalize(normal_xfm * n) (getNormals, common.h:233-264;
  // Pathtrace_impl.cpp:243-247; identity meshes: normalised twice)
  V3<R> surface_normal(int id, int prim) const {
    const Tri<R> &T = surface(id).tris[size_t(prim)];
    const V3<R> n = normalize(cross(T.e1, T.e2));
    if (id < int(meshes.size())) return normalize(n);
    return normalize(InstXf<R>::mul(ixf[size_t(id) - meshes.size()].N, n));
  }
  // intersectSurfaces: the closest (any: a) hit over the meshes, a later
  // mesh taking ties, then the instances, which must be strictly closer
  // (the lower instance index keeps ties)
  bool surfaces_hit(const Ray<R> &ray, bool any, R &t_out, int &id, int &prim, R &u,
                    R &v) const {
    bool hit = false;
    R best = ray.tmax;
    for (size_t m = 0; m < meshes.size(); ++m) {
      Ray<R> r = ray;
      if (hit) r.tmax = best;
      R t, uu, vv;
      int p = -1;
      if (meshes[m].intersect(r, any, t, p, uu, vv)) {
        hit = true;
        best = t;
        id = int(m);
        prim = p;
        u = uu;
        v = vv;
      }
    }
    for (size_t j = 0; j < imeshes.size(); ++j) {
      if (any && hit) break;
      Ray<R> r = ixf[j].to_object(ray);
      r.tmax = hit ? best : ray.tmax;
      R t, uu, vv;
      int p = -1;
      if (!imeshes[j].intersect(r, any, t, p, uu, vv)) continue;
      if (hit && !(t < best)) continue;
      hit = true;
      best = t;
      id = int(meshes.size() + j);
      prim = p;
      u = uu;
      v = vv;
    }
    t_out = best;
    return hit;
  }
};

template <class R> struct LightSample {
  V3<R> intensity, dir;
  R pdf = R(0), dist = R(0), dist2 = R(0);
};

// sampleLight (renderer/common.h:1432-1466) for directional lights;
// visionaray directional_light::sample with angularDiameter 0 returns the
// stored direction (Directional.cpp:33 stores -direction), pdf 1, intensity
// color*irradiance, distance FLT_MAX, and draws no random numbers
// (parity unpinned).
template <class R>
LightSample<R> sample_light(const orc_light &L, const QuadLight<R> &Q, const HdriLight &H,
                            V3<R> hitPos, Random &rnd) {
  LightSample<R> ls;
  if (L.type == ORC_LIGHT_HDRI) {
    // infinitely far, not attenuated (common.h:1448-1461)
    V3<float> d;
    float pdf;
    H.sample(rnd, d, pdf);
    ls.dir = V3<R>(R(d.x), R(d.y), R(d.z));
    ls.dist = R(std::numeric_limits<float>::max());
    const V3<float> I = H.radiance(d);
    ls.intensity = V3<R>(R(I.x), R(I.y), R(I.z));
    ls.pdf = R(pdf);
    ls.dist2 = R(1);
    return ls;
  }
  if (L.type == ORC_LIGHT_QUAD) {
    // area_light<float, Quad>::sample (visionaray, not vendored: parity
    // unpinned): sample_surface(Quad) picks a triangle with one draw
    // (DeviceCopyableObjects.h:2258-2270), then a uniform point with two
    // (sqrt parametrisation).  pdf = 1 / (A |cos theta_light|): with the
    // 1/dist^2 of sampleLight (common.h:1455-1463) the NEE estimate
    // f Le cos / (dist^2 pdf) is then the area-sampling estimator.
    const R r = R(rnd());
    const Tri<R> &t = (Q.A1 / (Q.A1 + Q.A2) < r) ? Q.t1 : Q.t2;
    const R u1 = R(rnd()), u2 = R(rnd());
    const R su = std::sqrt(u1);
    const V3<R> v2 = t.v1 + t.e1, v3 = t.v1 + t.e2;
    const V3<R> pos = t.v1 * (R(1) - su) + v2 * (su * (R(1) - u2)) + v3 * (su * u2);
    ls.dir = pos - hitPos;
    ls.dist = length(ls.dir);
    const R cosl = std::fabs(dot(Q.normal(), ls.dir * (R(1) / ls.dist)));
    ls.pdf = cosl > R(0) ? R(1) / ((Q.A1 + Q.A2) * cosl) : R(0);
    ls.intensity = Q.intensity;
    ls.dist2 = ls.dist * ls.dist;
    return ls;
  }
  V3<R> d = normalize(from<R>(L.dir));
  ls.dir = -d;
  ls.dist = R(std::numeric_limits<float>::max());
  ls.intensity = from<R>(L.color) * R(L.intensity);
  ls.pdf = R(1);
  ls.dist2 = R(1);  // infinite lights are not attenuated (common.h:1458-1461)
  return ls;
}

template <class R> struct HitRec {
  bool surf_hit = false;
  R surf_t = std::numeric_limits<R>::max();
  int surf_mesh = -1, surf_prim = -1;
  R surf_u = R(0), surf_v = R(0);
  int light_id = -1;
  VolHit<R> vol;
  int vol_inst = -1;
  bool light_hit = false;
  R light_t = std::numeric_limits<R>::max();
  bool light_visible = false;
  bool hit = false, volumeHit = false, lightHit = false;
};

template <class R>
HitRec<R> intersect_all(const Scene<R> &sc, const Ray<R> &ray, Random &rnd, Counter *cnt,
                        unsigned bounceID = 0, bool shadow = false, bool ratio = false) {
  HitRec<R> hr;
  // intersectSurfaces<1>: closest hit, any hit for shadow / AO rays
  {
    R t, u, v;
    int id = -1, prim = -1;
    if (sc.surfaces_hit(ray, shadow, t, id, prim, u, v)) {
      hr.surf_hit = true;
      hr.surf_t = t;
      hr.surf_mesh = id;
      hr.surf_prim = prim;
      hr.surf_u = u;
      hr.surf_v = v;
    }
  }
  // intersectLights (common.h:1566-1603): quads, skipping invisible lights
  // for camera rays; directional lights are never hit
  for (int i = 0; i < sc.s->num_lights; ++i) {
    const orc_light &L = sc.s->lights[i];
    if (bounceID == 0 && !L.visible) continue;
    if (L.type == ORC_LIGHT_HDRI) {
      // every ray meets the environment, at FLT_MAX (:1574-1588)
      if (!hr.light_hit) {
        hr.light_hit = true;
        hr.light_t = std::numeric_limits<R>::max();
        hr.light_visible = L.visible != 0;
        hr.light_id = i;
      }
      continue;
    }
    if (L.type != ORC_LIGHT_QUAD) continue;
    const TriHit<R> h = sc.quads[i].intersect(ray);
    if (h.hit && h.t < hr.light_t) {
      hr.light_hit = true;
      hr.light_t = h.t;
      hr.light_visible = L.visible != 0;
      hr.light_id = i;
    }
  }
  // sampleFreeFlightDistanceAllVolumes -> intersectVolumes, closest wins;
  // ratio tracking (shadow / AO rays) multiplies the volumes' estimates
  for (size_t i = 0; i < sc.vols.size(); ++i) {
    Ray<R> lr = sc.vols[i].to_local(ray);
    VolHit<R> vh = delta_track(sc.vols[i], lr, rnd, cnt, ratio);
    if (ratio) {
      hr.vol.Tr = hr.vol.Tr * vh.Tr;
    } else if (vh.hit && vh.t < hr.vol.t) {
      hr.vol = vh;
      hr.vol_inst = int(i);
    }
  }
  hr.hit = hr.surf_hit || hr.vol.hit || hr.light_hit;
  hr.lightHit = hr.light_hit && (!hr.surf_hit || hr.light_t <= hr.surf_t) &&
                (!hr.vol.hit || hr.light_t <= hr.vol.t);
  hr.volumeHit = hr.vol.hit && (!hr.surf_hit || hr.vol.t < hr.surf_t) &&
                 (!hr.light_hit || hr.vol.t < hr.light_t);
  return hr;
}

// A4: clipRay (renderer/common.h:1469-1482)
template <class R> void clip_ray(Ray<R> &ray, const orc_params *P) {
  for (int i = 0; i < P->num_clip_planes; ++i) {
    const float *c = P->clip_planes[i];
    const V3<R> n{R(c[0]), R(c[1]), R(c[2])};
    const R s = dot(n, ray.dir);
    if (s != R(0)) {
      const R t = (R(c[3]) - dot(n, ray.ori)) / s;
      if (s < R(0))
        ray.tmin = std::max(ray.tmin, t);
      else
        ray.tmax = std::min(ray.tmax, t);
    }
  }
}

// ---------------------------------------------------------------------------
// A16: path tracer (renderer/Pathtrace_impl.cpp)
// ---------------------------------------------------------------------------
enum RayType { Radiance, Shadow, AO, Miss, None };

// the PixelSample fields of the g-buffer (frame/common.h:15-26); normal and
// albedo, unset there for a sample without a hit, are 0 here
template <class R> struct GBuf {
  R depth = R(1e31f);
  V3<R> Ng = V3<R>(R(0)), albedo = V3<R>(R(0));
  uint32_t prim = ~0u, obj = ~0u, inst = ~0u;
  // TAA: motion vector of the sample's last radiance vertex, with the
  // previous / current proj * view matrices it is measured with
  float mvx = 0.f, mvy = 0.f;
  const float *vp_prev = nullptr, *vp_curr = nullptr;
  int W = 0, H = 0;
};

// project (visionaray, parity unpinned: gluProject) into window coordinates
inline void project_window(const float *M, V3<float> p, int W, int H, float &wx, float &wy) {
  const float x = ((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3];
  const float y = ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7];
  const float w = ((M[12] * p.x + M[13] * p.y) + M[14] * p.z) + M[15];
  wx = float(W) * (x / w + 1.f) / 2.f;
  wy = float(H) * (y / w + 1.f) / 2.f;
}

// the motion vector of a radiance vertex (Pathtrace_impl.cpp:290-297:
// every vertex overwrites it)
template <class R> inline void gbuf_motion(GBuf<R> *gb, V3<R> hitPos) {
  if (!gb || !gb->vp_curr) return;
  const V3<float> p(float(hitPos.x), float(hitPos.y), float(hitPos.z));
  float ax, ay, bx, by;
  project_window(gb->vp_prev, p, gb->W, gb->H, ax, ay);
  project_window(gb->vp_curr, p, gb->W, gb->H, bx, by);
  gb->mvx = ax - bx;
  gb->mvy = ay - by;
}

template <class R> struct ShadeState {
  V3<R> baseColor, shadedColor, hitPosOff, sn;
  int aoSamples = 0;
  R aoWeights = 0, aoCount = 0;
  // bsdfSample
  V3<R> bs_dir, bs_f;
  R bs_pdf = 0, bs_cosT = 0;
  V3<R> bs_w;  // path weight f cos / pdf of the sample
  LightSample<R> ls;
  R misWeightNEE = 0;
  RayType nextType = None;
  Ray<R> next;
  R vis_light = 1, vis_ao = 1;
};

// power_heuristic (visionaray, parity unpinned): f^2 / (f^2 + g^2), 0 when
// both vanish
template <class R> inline R power_heuristic(R f, R g) {
  const R d = f * f + g * g;
  return d > R(0) ? (f * f) / d : R(0);
}

template <class R>
void shade(Random &srnd, const Ray<R> &ray, RayType rayType, const Scene<R> &sc,
           const orc_params *P, const HitRec<R> &hitRec, ShadeState<R> &st,
           unsigned bounceID = 0, GBuf<R> *gb = nullptr) {
  st.nextType = Miss;
  if (rayType == Radiance) {
    st.baseColor = V3<R>(R(0));
    st.vis_light = R(1);
    st.vis_ao = R(1);
    st.aoSamples = 0;
    st.aoWeights = R(0);
    st.aoCount = R(0);
    if (!hitRec.hit) {
      st.shadedColor = V3<R>(R(0));
      st.nextType = Miss;
      return;
    }
    const int numLights = sc.s->num_lights;
    if (hitRec.lightHit && sc.s->lights[hitRec.light_id].type == ORC_LIGHT_HDRI) {
      // Pathtrace_impl.cpp:144-182: the map in the ray's direction,
      // MIS-weighted after the first bounce
      const HdriLight &H = sc.hdris[hitRec.light_id];
      const V3<float> d(float(ray.dir.x), float(ray.dir.y), float(ray.dir.z));
      const V3<float> I = H.radiance(d);
      st.shadedColor = V3<R>(R(I.x), R(I.y), R(I.z));
      st.misWeightNEE = R(1);
      if (bounceID > 0)
        st.shadedColor =
            st.shadedColor * power_heuristic(st.bs_pdf, R(H.pdf(d)) / R(numLights));
      st.nextType = Miss;
      return;
    }
    if (hitRec.lightHit) {
      // Pathtrace_impl.cpp:139-183: the quad's intensity, MIS-weighted
      // against the BSDF sample after the first bounce
      const QuadLight<R> &Q = sc.quads[hitRec.light_id];
      const V3<R> hitPos = ray.ori + ray.dir * hitRec.light_t;
      st.shadedColor = Q.intensity;
      st.misWeightNEE = R(1);
      if (bounceID > 0) {
        const R A = Q.A1 + Q.A2;
        const R ld = length(hitPos - ray.ori);
        const V3<R> L = normalize(hitPos - ray.ori);
        const R LdotNl = std::fabs(dot(-L, Q.normal()));
        const R solidAngle = (LdotNl * A) / (ld * ld);
        const R lightPDF = R(1) / solidAngle;
        st.shadedColor = st.shadedColor * power_heuristic(st.bs_pdf, lightPDF / R(numLights));
      }
      st.nextType = Miss;
      return;
    }
    V3<R> hitPos, gn, sn, color;
    R eps;
    V3<R> viewDir = -normalize(ray.dir);
    if (hitRec.volumeHit) {
      // Pathtrace_impl.cpp:203-233
      hitPos = ray.ori + ray.dir * hitRec.vol.t;
      eps = epsilon_from(hitPos, ray.dir, hitRec.vol.t);
      gn = V3<R>(R(0));
      const Volume<R> &V = sc.vols[hitRec.vol_inst];
      if (P->gradient_shading) {  // :204-212, at the hit in voxel space
        const Ray<R> lr = V.to_local(ray);
        gn = V.gradient_normal(V.point_to_voxel(lr.ori + lr.dir * hitRec.vol.t));
      }
      if (P->ambient_samples > 0 && length(gn) < R(1e-3)) {
        R u1 = R(srnd()), u2 = R(srnd());
        gn = uniform_sample_sphere(u1, u2);
      }
      sn = gn;
      // getNormalTransform (renderer/common.h:766-770): transpose(affineInv),
      // no renormalisation for volumes (:217-219)
      gn = V.mul_normal(gn);
      sn = V.mul_normal(sn);
      sn = faceforward(sn, viewDir, gn);
      gn = faceforward(gn, viewDir, gn);
      color = hitRec.vol.albedo;
      if (gb && bounceID == 0) {  // :228-233
        gb->depth = hitRec.vol.t;
        gb->prim = uint32_t(hitRec.vol.prim);
        gb->obj = V.v->object_id;
        gb->inst = V.v->instance_id;
      }
    } else {
      // triangle surface (Pathtrace_impl.cpp:234-274): Ng = normalize(e1 x
      // e2), no vertex normals (getNormals, common.h:233-264), matte color
      const Mesh<R> &M = sc.surface(hitRec.surf_mesh);
      hitPos = ray.ori + ray.dir * hitRec.surf_t;
      eps = epsilon_from(hitPos, ray.dir, hitRec.surf_t);
      // the instance's normal transform, then normalised again (:243-247)
      gn = sc.surface_normal(hitRec.surf_mesh, hitRec.surf_prim);
      sn = gn;
      sn = faceforward(sn, viewDir, gn);
      gn = faceforward(gn, viewDir, gn);
      color = M.color;
      if (gb && bounceID == 0) {  // :273-278
        gb->depth = hitRec.surf_t;
        gb->prim = uint32_t(hitRec.surf_prim);
        gb->obj = M.m->object_id;
        gb->inst = M.m->instance_id;
      }
    }
    if (gb && bounceID == 0) {  // :281-285
      gb->Ng = gn;
      gb->albedo = color;
    }
    gbuf_motion(gb, hitPos);
    st.sn = sn;
    st.hitPosOff = hitPos + sn * eps;

    int lightID = -1;
    if (numLights > 0) {
      // uniformSampleOneLight, renderer/common.h:165-169
      int which = int(srnd() * float(numLights));
      if (which == numLights) which = 0;
      lightID = which;
      st.ls = sample_light<R>(sc.s->lights[lightID], sc.quads[lightID], sc.hdris[lightID], hitPos,
                              srnd);
    }
    auto safe_rcp = [](R f) { return f > R(0) ? R(1) / f : R(0); };
    V3<R> lightIntensity = st.ls.intensity * safe_rcp(st.ls.dist2);
    if (!hitRec.volumeHit) {
      // matte (common.h:1033-1048, 1371-1413): f = color / pi, cosine
      // hemisphere around faceforward(Ns, viewDir, Ng), pdf = cos / pi
      const R inv_pi = R(0.318309886183790671538);
      const V3<R> f = color * inv_pi;
      const V3<R> lightDir = normalize(st.ls.dir);
      const R NdotL = std::max(R(0), dot(sn, lightDir));
      st.shadedColor = f * lightIntensity * NdotL * safe_rcp(st.ls.pdf) * R(numLights);
      const V3<R> w = faceforward(sn, viewDir, gn);
      V3<R> u, v;
      make_onb(w, u, v);
      const R u1 = R(srnd()), u2 = R(srnd());
      const V3<R> sp = cosine_sample_hemisphere(u1, u2);
      st.bs_dir = normalize(u * sp.x + v * sp.y + w * sp.z);
      st.bs_pdf = std::max(R(0), dot(sn, st.bs_dir)) * inv_pi;
      st.bs_cosT = std::max(R(0), dot(sn, st.bs_dir));
      st.bs_f = f;
      st.bs_w = f * (st.bs_cosT * safe_rcp(st.bs_pdf));
    }
    if (hitRec.volumeHit) {
      // gradient shading: matte NEE along the gradient normal (:311-328);
      // else :330-339, no phase factor, no light count
      const bool matte = P->gradient_shading && length(gn) > R(1e-10);
      if (matte) {
        const R inv_pi = R(0.318309886183790671538);
        const R NdotL = std::max(R(0), dot(sn, normalize(st.ls.dir)));
        st.shadedColor = hitRec.vol.albedo * inv_pi * lightIntensity * NdotL *
                         safe_rcp(st.ls.pdf) * R(numLights);
      } else {
        st.shadedColor = hitRec.vol.albedo * lightIntensity * safe_rcp(st.ls.pdf) *
                         safe_rcp(st.ls.dist2);
      }
      R u1 = R(srnd()), u2 = R(srnd());
      if (P->phase_g == 0.f) {
        st.bs_dir = uniform_sample_sphere(u1, u2);
        st.bs_pdf = R(1);
      } else {
        // anisotropic phase (an extension): NEE carries 4 pi p_HG towards
        // the light, the sample is drawn from p_HG and weighs the albedo
        const R g = R(P->phase_g);
        const V3<R> d = normalize(ray.dir);
        if (numLights > 0 && !matte)
          st.shadedColor = st.shadedColor * hg_4pi(g, dot(d, normalize(st.ls.dir)));
        st.bs_dir = hg_sample(g, d, u1, u2);
        st.bs_pdf = hg_4pi(g, dot(d, st.bs_dir));
      }
      st.bs_f = hitRec.vol.albedo;
      st.bs_cosT = R(1);
      st.bs_w = hitRec.vol.albedo * (R(1) * R(1));
    }
    st.misWeightNEE = R(0);
    if (numLights > 0) {
      // MIS against the BSDF sample for area lights, 1 for delta lights
      // (:370-380)
      if (sc.s->lights[lightID].type == ORC_LIGHT_QUAD ||
          sc.s->lights[lightID].type == ORC_LIGHT_HDRI)
        st.misWeightNEE = power_heuristic(st.ls.pdf / R(numLights), st.bs_pdf);
      else
        st.misWeightNEE = R(1);
    }
    st.baseColor = color;
  }

  // state machine, Pathtrace_impl.cpp:401-486
  auto prepShadow = [&]() {
    st.next.ori = st.hitPosOff;
    st.next.dir = normalize(st.ls.dir);
    st.next.tmin = R(0);
    st.next.tmax = st.ls.dist;
    st.nextType = Shadow;
  };
  auto prepAO = [&]() {
    V3<R> u, v, w = st.sn;
    make_onb(w, u, v);
    R u1 = R(srnd()), u2 = R(srnd());
    V3<R> sp = cosine_sample_hemisphere(u1, u2);
    st.next.ori = st.hitPosOff;
    st.next.dir = normalize(u * sp.x + v * sp.y + w * sp.z);
    st.next.tmin = R(0);
    st.next.tmax = R(P->occlusion_distance);
    st.nextType = AO;
  };
  auto prepRadiance = [&]() {
    st.next.ori = st.hitPosOff;
    st.next.dir = normalize(st.bs_dir);
    st.next.tmin = R(0);
    st.next.tmax = R(1e31f);
    st.nextType = Radiance;
  };

  if (rayType == Radiance) {
    if (st.ls.pdf >= R(0)) { prepShadow(); return; }
    if (st.aoSamples < P->ambient_samples) { prepAO(); return; }
    if (st.bs_pdf >= R(0)) { prepRadiance(); return; }
    return;
  } else if (rayType == Shadow) {
    int surfV = hitRec.surf_hit ? 0 : 1;
    int volV = hitRec.volumeHit ? 0 : 1;
    st.vis_light = R(surfV * volV) * hitRec.vol.Tr;
    if (st.aoSamples < P->ambient_samples) { prepAO(); return; }
    if (st.bs_pdf >= R(0)) { prepRadiance(); return; }
    return;
  } else if (rayType == AO) {
    st.aoSamples++;
    R weight = std::max(R(0), dot(ray.dir, st.sn));
    st.aoWeights += weight;
    // occluded by any hit; with ratio tracking a volume occludes by 1 - Tr
    // (delta tracking: Tr is 0 at a volume collision, else 1)
    if (weight > R(0)) st.aoCount += weight * (hitRec.hit ? R(1) : R(1) - hitRec.vol.Tr);
    if (st.aoSamples < P->ambient_samples) { prepAO(); return; }
    R aoV = R(0);
    if (st.aoWeights > R(0)) aoV = R(1) - st.aoCount / st.aoWeights;
    st.vis_ao *= aoV;
    if (st.bs_pdf >= R(0)) { prepRadiance(); return; }
    return;
  }
}

template <class R>
V4<R> pathtrace_pixel(const Scene<R> &sc, const Camera<R> &cam, const orc_params *P,
                      int x, int y, int W, int H, int accumID, Counter *cnt,
                      GBuf<R> *closest = nullptr) {
  // Pathtrace_impl.cpp:510-513: rng seeded by (pixelID, accumID); the
  // ScreenSample holds a copy, RR draws from the un-advanced original
  uint32_t pixelID = uint32_t(x + W * y);
  Random rng(pixelID, uint32_t(accumID) + P->rng_offset);
  Random srnd = rng;
  V4<R> accum{0, 0, 0, 0};
  const int spp = P->pixel_samples;
  GBuf<R> proto;  // the TAA matrices, carried into every sample's record
  if (closest) {
    proto.vp_prev = closest->vp_prev;
    proto.vp_curr = closest->vp_curr;
    proto.W = closest->W;
    proto.H = closest->H;
    *closest = proto;
  }
  for (int sampleID = 0; sampleID < spp; ++sampleID) {
    GBuf<R> ps = proto;
    R xf = R(x), yf = R(y);
    R jx = R(srnd()) - R(0.5);
    R jy = R(srnd()) - R(0.5);
    xf += jx;
    yf += jy;
    Ray<R> ray = cam.primary(xf, yf, R(W), R(H));
    if (cam.lens_radius > R(0)) {  // the lens sample precedes the shutter draw
      const R u1 = R(srnd()), u2 = R(srnd());
      thin_lens(cam, ray, u1, u2);
    }
    (void)srnd();  // shutter time lerp, DeviceCopyableObjects.h:2597
    V4<R> color{R(P->bg[0]), R(P->bg[1]), R(P->bg[2]), R(P->bg[3])};
    if (!sc.empty()) {
      bool firstHit = false, firstLightInvisible = false;
      ShadeState<R> st;
      RayType rayType = Radiance;
      V3<R> throughput(R(1)), intensity(R(0));
      for (unsigned passID = 0, bounceID = 0;; ++passID) {
        clip_ray(ray, P);  // Pathtrace_impl.cpp:559
        const bool shadow = rayType == Shadow || rayType == AO;
        HitRec<R> hitRec = intersect_all(sc, ray, srnd, cnt, bounceID, shadow,
                                         shadow && P->ratio_tracking != 0);
        shade(srnd, ray, rayType, sc, P, hitRec, st, bounceID, closest ? &ps : nullptr);
        if (passID == 0 && bounceID == 0) {
          firstHit = hitRec.hit;
          firstLightInvisible = hitRec.lightHit && !hitRec.light_visible;
        }
        ray = st.next;
        rayType = st.nextType;
        if (rayType == Miss || rayType == Radiance) {
          V3<R> direct = st.shadedColor * st.vis_light;
          V3<R> ambient = st.baseColor * from<R>(P->ambient_color) *
                          R(P->ambient_radiance) * st.vis_ao;
          intensity = intensity + throughput * (direct * st.misWeightNEE);
          intensity = intensity + throughput * ambient;
          throughput = throughput * st.bs_w;
          bounceID++;
        }
        if (rayType == Miss || int(bounceID) > P->max_bounce) break;
        if (bounceID > 1 && g_roulette) {  // Russian roulette, :594-599
          R prob = max_element(throughput);
          if (R(rng()) > prob) break;
          throughput = throughput * (R(1) / prob);
        }
      }
      if (firstHit && !firstLightInvisible) color = V4<R>{intensity.x, intensity.y, intensity.z, R(1)};
    }
    accum.x += color.x;
    accum.y += color.y;
    accum.z += color.z;
    accum.w += color.w;
    // the closest sample's g-buffer, the first of equal depths (:616-618)
    if (closest && ps.depth < closest->depth) *closest = ps;
  }
  R inv = R(1) / R(spp);
  return V4<R>{accum.x * inv, accum.y * inv, accum.z * inv, accum.w * inv};
}

// ---------------------------------------------------------------------------
// A14/A15: raycast (Raycast_impl.cpp, VolumeIntegration.h:374-470)
// ---------------------------------------------------------------------------
// Gradient shading of one ray-caster sample (VolumeIntegration.h:420-459):
// matte under every light along the faceforwarded gradient normal at the
// voxel-space point Pv (lights sampled there, as the reference does) plus
// the ambient term
template <class R>
V3<R> gradient_shade(const Scene<R> &sc, const Volume<R> &vol, const orc_params *P, V3<R> Pv,
                     V3<R> viewDir, V3<R> albedo, Random &rs) {
  V3<R> gn = vol.gradient_normal(Pv);
  gn = faceforward(gn, viewDir, gn);
  V3<R> shaded(R(0));
  auto safe_rcp = [](R f) { return f > R(0) ? R(1) / f : R(0); };
  if (length(gn) > R(1e-10)) {
    const R inv_pi = R(0.318309886183790671538);
    for (int i = 0; i < sc.s->num_lights; ++i) {
      const LightSample<R> ls = sample_light<R>(sc.s->lights[i], sc.quads[i], sc.hdris[i], Pv, rs);
      const V3<R> lightDir = normalize(ls.dir);
      const V3<R> lightIntensity = ls.intensity * safe_rcp(ls.dist2);
      const R NdotL = std::max(R(0), dot(gn, lightDir));
      shaded = shaded + albedo * inv_pi * lightIntensity * NdotL * safe_rcp(ls.pdf);
    }
  }
  return shaded + albedo * from<R>(P->ambient_color) * R(P->ambient_radiance);
}

template <class R>
V4<R> raycast_pixel(const Scene<R> &sc, const Camera<R> &cam, const orc_params *P,
                    int x, int y, int W, int H, GBuf<R> *gb = nullptr) {
  V4<R> bg{R(P->bg[0]), R(P->bg[1]), R(P->bg[2]), R(P->bg[3])};
  if (gb) *gb = GBuf<R>();
  if (sc.empty()) return bg;
  // light samples (surface and gradient shading) come from the pixel's
  // path-tracer stream: the reference's ray caster leaves its generator
  // unseeded; a thin lens's sample, if any, comes first
  Random rs(uint32_t(x + W * y), 0u);
  Ray<R> ray = cam.primary(R(x), R(y), R(W), R(H));
  if (cam.lens_radius > R(0)) {
    const R u1 = R(rs()), u2 = R(rs());
    thin_lens(cam, ray, u1, u2);
  }
  clip_ray(ray, P);  // Raycast_impl.cpp:288
  auto safe_rcp = [](R f) { return f > R(0) ? R(1) / f : R(0); };
  // the closest surface (Raycast_impl.cpp:27-156): matte under every light
  // along the surface normal (not faceforwarded) plus the ambient term;
  // matte is opaque, so the first hit ends the loop and bounds the volume
  V4<R> surf{R(0), R(0), R(0), R(0)};
  {
    R st = std::numeric_limits<R>::max();
    int sm = -1, sp = -1;
    {
      R t, u, v;
      if (sc.surfaces_hit(ray, false, t, sm, sp, u, v))
        st = t;
      else
        sm = -1;
    }
    if (sm >= 0) {
      const Mesh<R> &M = sc.surface(sm);
      const V3<R> hitPos = ray.ori + ray.dir * st;
      const V3<R> sn = sc.surface_normal(sm, sp);
      const R inv_pi = R(0.318309886183790671538);
      V3<R> shaded(R(0));
      for (int i = 0; i < sc.s->num_lights; ++i) {
        const LightSample<R> ls =
            sample_light<R>(sc.s->lights[i], sc.quads[i], sc.hdris[i], hitPos, rs);
        const V3<R> lightDir = normalize(ls.dir);
        const V3<R> lightIntensity = ls.intensity * safe_rcp(ls.dist2);
        const R NdotL = std::max(R(0), dot(sn, lightDir));
        shaded = shaded + M.color * inv_pi * lightIntensity * NdotL * safe_rcp(ls.pdf);
      }
      shaded = shaded + M.color * from<R>(P->ambient_color) * R(P->ambient_radiance);
      surf = V4<R>{shaded.x, shaded.y, shaded.z, R(1)};
      ray.tmax = st;
      if (gb) {  // :135-143
        gb->depth = st;
        gb->Ng = sn;
        gb->albedo = M.color;
        gb->prim = uint32_t(sp);
        gb->obj = M.m->object_id;
        gb->inst = M.m->instance_id;
      }
    }
  }
  // over(float4(surfaceColor, surfaceAlpha), backdrop), Raycast_impl.cpp:
  // 158-173: the backdrop is the closest visible light the ray meets -- a
  // quad, or the environment at FLT_MAX -- as float4(intensity, 1), else the
  // background colour
  V4<R> result = bg;
  {
    int lid = -1;
    R lt = std::numeric_limits<R>::max();
    for (int i = 0; i < sc.s->num_lights; ++i) {
      const orc_light &L = sc.s->lights[i];
      if (!L.visible) continue;
      if (L.type == ORC_LIGHT_HDRI) {
        if (lid < 0) lid = i;
        continue;
      }
      if (L.type != ORC_LIGHT_QUAD) continue;
      const TriHit<R> h = sc.quads[i].intersect(ray);
      if (h.hit && h.t < lt) {
        lt = h.t;
        lid = i;
      }
    }
    if (lid >= 0) {
      if (sc.s->lights[lid].type == ORC_LIGHT_HDRI) {
        const V3<float> I = sc.hdris[lid].radiance(
            V3<float>(float(ray.dir.x), float(ray.dir.y), float(ray.dir.z)));
        result = V4<R>{R(I.x), R(I.y), R(I.z), R(1)};
      } else {
        const V3<R> I = sc.quads[lid].intensity;
        result = V4<R>{I.x, I.y, I.z, R(1)};
      }
    }
    const R a = R(1) - surf.w;
    result = V4<R>{surf.x + a * result.x, surf.y + a * result.y, surf.z + a * result.z,
                   surf.w + a * result.w};
  }
  // intersectVolumeBounds: closest volume box (DeviceCopyableObjects.h:695-709)
  int best = -1;
  R bestT = std::numeric_limits<R>::max();
  for (size_t i = 0; i < sc.vols.size(); ++i) {
    Ray<R> lr = sc.vols[i].to_local(ray);
    BoxHit<R> bh = intersect_box(lr, sc.vols[i].blo, sc.vols[i].bhi);
    if (!bh.hit || !(bh.tfar >= lr.tmin)) continue;  // :695-700
    R t = std::max(lr.tmin, bh.tnear);
    if (t < bestT) {
      bestT = t;
      best = int(i);
    }
  }
  if (best < 0) return result;
  const Volume<R> &vol = sc.vols[best];
  if (gb) {  // :230-237: the volume's ids; normal and albedo 0
    gb->obj = vol.v->object_id;
    gb->inst = vol.v->instance_id;
    gb->Ng = V3<R>(R(0));
    gb->albedo = V3<R>(R(0));
  }
  Ray<R> lr = vol.to_local(ray);
  BoxHit<R> bh = intersect_box(lr, vol.blo, vol.bhi);
  lr.tmin = std::max(lr.tmin, bh.tnear);
  lr.tmax = std::min(lr.tmax, bh.tfar);
  V3<R> ori = vol.point_to_voxel(lr.ori);
  V3<R> dir = vol.vector_to_voxel(lr.dir);  // not renormalised
  const R samplingRateInv = P->volume_sampling_rate > 0.f ? R(1) / R(P->volume_sampling_rate) : R(0);
  const R dt = vol.cellSize * samplingRateInv;
  const R expo = dt / R(vol.v->unit_distance);
  V3<R> color(R(0));
  R alpha = R(0), transmittance = R(1);
  for (R t = lr.tmin; t <= lr.tmax && alpha < R(0.99f); t += dt) {
    V3<R> p = ori + dir * t;
    R v;
    if (!vol.sample(p, v)) continue;  // VolumeIntegration.h:412
    V4<R> s = vol.classify(v);
    R stepT = opacity_pow(R(1) - s.w, expo);
    V3<R> c(s.x, s.y, s.z);
    if (P->gradient_shading) c = gradient_shade(sc, vol, P, p, -dir, c, rs);
    color = color + c * (transmittance * (R(1) - stepT));
    alpha += transmittance * (R(1) - stepT);
    transmittance *= stepT;
  }
  // over(alpha*float4(color,1), result) -- Raycast_impl.cpp:232
  V4<R> A{alpha * color.x, alpha * color.y, alpha * color.z, alpha};
  R k = R(1) - A.w;
  return V4<R>{A.x + k * result.x, A.y + k * result.y, A.z + k * result.z, A.w + k * result.w};
}

// ---------------------------------------------------------------------------
// thread pool over 64x64 tiles (for_each.h:76-84)
// ---------------------------------------------------------------------------
template <class F>
void parallel_tiles(int W, int H, int threads, F &&fn) {
  const int T = 64;
  int tx = (W + T - 1) / T, ty = (H + T - 1) / T;
  int ntiles = tx * ty;
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  std::atomic<int> next(0);
  auto work = [&]() {
    while (true) {
      int t = next.fetch_add(1);
      if (t >= ntiles) break;
      int bx = (t % tx) * T, by = (t / tx) * T;
      for (int y = by; y < std::min(by + T, H); ++y)
        for (int x = bx; x < std::min(bx + T, W); ++x) fn(x, y);
    }
  };
  if (threads == 1) {
    work();
    return;
  }
  std::vector<std::thread> pool;
  for (int i = 0; i < threads; ++i) pool.emplace_back(work);
  for (auto &th : pool) th.join();
}

}  // namespace orc

using namespace orc;

// ===========================================================================
// C API
// ===========================================================================
namespace orc {

// max TF opacity over the taps tex1D reads for values in [vmin, vmax],
// quantized like the GPU's 8-bit majorants
inline float tf_majorant(const orc_volume *v, float vmin, float vmax) {
  const int n = v->tf_size;
  float u0 = (vmin - v->value_lo) / (v->value_hi - v->value_lo);
  float u1 = (vmax - v->value_lo) / (v->value_hi - v->value_lo);
  float c0 = std::min(std::max(u0 * n - 0.5f, -1.f), float(n));
  float c1 = std::min(std::max(u1 * n - 0.5f, -1.f), float(n));
  int lo = std::min(std::max(int(std::floor(c0)), 0), n - 1);
  int hi = std::min(std::max(int(std::floor(c1)) + 1, 0), n - 1);
  float m = 0.f;
  for (int i = lo; i <= hi; ++i) m = std::max(m, v->tf_rgba[4 * i + 3]);
  return quantize_majorant(m, majorant_step(v));
}

// nanovdb majorants: a cell of the grid over the float world bounds reads
// voxels in the box from the index images of its 8 corners (double map),
// floor, one voxel below and two above; the value range is the union over
// the 8^3 slots (leaf-aligned blocks) that box touches
void build_majorants_nanovdb(const orc_volume *v, const int G[3], float *out) {
  NanoVDBGrid g;
  if (!g.open(v->nanovdb)) return;
  float glo[3], ghi[3];
  for (int a = 0; a < 3; ++a) {
    glo[a] = float(g.wlo[a]);
    ghi[a] = float(g.whi[a]);
  }
  std::map<std::array<int, 3>, std::pair<float, float>> slot_range;
  auto range_of = [&](int sx, int sy, int sz) {
    auto key = std::array<int, 3>{sx, sy, sz};
    auto it = slot_range.find(key);
    if (it != slot_range.end()) return it->second;
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 8; ++j)
        for (int k = 0; k < 8; ++k) {
          const float val = g.get(sx * 8 + i, sy * 8 + j, sz * 8 + k);
          mn = std::min(mn, val);
          mx = std::max(mx, val);
        }
    slot_range.emplace(key, std::make_pair(mn, mx));
    return std::make_pair(mn, mx);
  };
  for (int cz = 0; cz < G[2]; ++cz)
    for (int cy = 0; cy < G[1]; ++cy)
      for (int cx = 0; cx < G[0]; ++cx) {
        const int c[3] = {cx, cy, cz};
        double wmin[3], wmax[3];
        for (int a = 0; a < 3; ++a) {
          const double w = (double(ghi[a]) - double(glo[a])) / double(G[a]);
          wmin[a] = double(glo[a]) + double(c[a]) * w;
          wmax[a] = double(glo[a]) + double(c[a] + 1) * w;
        }
        double imin[3] = {INFINITY, INFINITY, INFINITY}, imax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 8; ++k) {
          const double p[3] = {(k & 1 ? wmax[0] : wmin[0]) - g.vec_d[0],
                               (k & 2 ? wmax[1] : wmin[1]) - g.vec_d[1],
                               (k & 4 ? wmax[2] : wmin[2]) - g.vec_d[2]};
          for (int a = 0; a < 3; ++a) {
            const double i = g.inv_d[3 * a] * p[0] + g.inv_d[3 * a + 1] * p[1] + g.inv_d[3 * a + 2] * p[2];
            imin[a] = std::min(imin[a], i);
            imax[a] = std::max(imax[a], i);
          }
        }
        int s0[3], s1[3];
        for (int a = 0; a < 3; ++a) {
          s0[a] = (int(std::floor(imin[a])) - 1) >> 3;
          s1[a] = (int(std::floor(imax[a])) + 2) >> 3;
        }
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (int sz = s0[2]; sz <= s1[2]; ++sz)
          for (int sy = s0[1]; sy <= s1[1]; ++sy)
            for (int sx = s0[0]; sx <= s1[0]; ++sx) {
              const auto r = range_of(sx, sy, sz);
              mn = std::min(mn, r.first);
              mx = std::max(mx, r.second);
            }
        out[(size_t(cz) * G[1] + cy) * G[0] + cx] = tf_majorant(v, mn, mx);
      }
}

// unstructured majorants: buildGrid (UnstructuredField.cpp:376-399) ->
// rasterizeBox (GridAccel.h:51-67): every cell's box, with its value range,
// into the grid cells from projectOnGrid (GridAccel-common.h:18-24; the
// float image is clamped to [-1, G] before the int conversion, which leaves
// in-range results as they are); cells no box reaches get 0
void build_majorants_unstructured(const orc_volume *v, const int G[3], float *out) {
  TetField f;
  const size_t n = size_t(G[0]) * G[1] * G[2];
  if (!f.open(v)) {
    std::fill(out, out + n, 0.f);
    return;
  }
  std::vector<float> mn(n, FLT_MAX), mx(n, -FLT_MAX);
  auto project = [&](float p, int a) {
    const float c = std::min(std::max((p - f.blo[a]) / (f.bhi[a] - f.blo[a]) * float(G[a]), -1.f),
                             float(G[a]));
    return std::min(std::max(int(c), 0), G[a] - 1);
  };
  for (const auto &c : f.cells) {
    int c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
      c0[a] = project(c.lo[a], a);
      c1[a] = project(c.hi[a], a);
    }
    for (int z = c0[2]; z <= c1[2]; ++z)
      for (int y = c0[1]; y <= c1[1]; ++y)
        for (int x = c0[0]; x <= c1[0]; ++x) {
          const size_t i = (size_t(z) * G[1] + y) * G[0] + x;
          mn[i] = std::min(mn[i], c.vmin);
          mx[i] = std::max(mx[i], c.vmax);
        }
  }
  for (size_t i = 0; i < n; ++i) out[i] = mx[i] < mn[i] ? 0.f : tf_majorant(v, mn[i], mx[i]);
}

}  // namespace orc

extern "C" {

uint32_t orc_tea_seed(uint32_t v0, uint32_t v1) { return Random(v0, v1).state; }

void orc_lcg_stream(uint32_t s0, uint32_t s1, int n, float *out) {
  Random r(s0, s1);
  for (int i = 0; i < n; ++i) out[i] = r();
}

int orc_tf_build(const float *colors, int numColors, int nch, const float *opac,
                 int numOpac, const float uni[4], float *out) {
  // TransferFunction1D.cpp:74-111
  size_t nC = colors ? size_t(numColors) : 1;
  size_t nO = opac ? size_t(numOpac) : 1;
  size_t tfSize = std::max(nC, nO);
  if (!out) return int(tfSize);
  for (size_t i = 0; i < tfSize; ++i) {
    float colorPos = tfSize > 1 ? (float(i) / float(tfSize - 1)) * float(nC - 1) : 0.f;
    float colorFrac = colorPos - std::floor(colorPos);
    float c0[4] = {uni[0], uni[1], uni[2], uni[3]};
    float c1[4] = {uni[0], uni[1], uni[2], uni[3]};
    if (colors) {
      int a = int(std::floor(colorPos)), b = int(std::ceil(colorPos));
      for (int k = 0; k < nch; ++k) {
        c0[k] = colors[a * nch + k];
        c1[k] = colors[b * nch + k];
      }
    }
    float c[4];
    for (int k = 0; k < 4; ++k) c[k] = c0[k] + colorFrac * (c1[k] - c0[k]);  // lerp_r
    if (opac) {
      float alphaPos = tfSize > 1 ? (float(i) / float(tfSize - 1)) * float(nO - 1) : 0.f;
      float alphaFrac = alphaPos - std::floor(alphaPos);
      float a0 = opac[int(std::floor(alphaPos))], a1 = opac[int(std::ceil(alphaPos))];
      c[3] *= a0 + alphaFrac * (a1 - a0);
    }
    for (int k = 0; k < 4; ++k) out[4 * i + k] = c[k];
  }
  return int(tfSize);
}

void orc_sample_field(const orc_volume *v, int n, const float *p, float *out) {
  Volume<float> vol(v);
  for (int i = 0; i < n; ++i) {
    float val;
    out[i] = vol.sample(V3<float>(p[3 * i], p[3 * i + 1], p[3 * i + 2]), val)
                 ? val
                 : std::numeric_limits<float>::quiet_NaN();
  }
}

void orc_sample_field_brute(const orc_volume *v, int n, const float *p, float *out) {
  orc::TetField f;
  const bool ok = v->field_type == 2 && f.open(v);
  for (int i = 0; i < n; ++i) {
    float val;
    out[i] = ok && f.sample_brute(V3<float>(p[3 * i], p[3 * i + 1], p[3 * i + 2]), val)
                 ? val
                 : std::numeric_limits<float>::quiet_NaN();
  }
}

int orc_unstructured_info(const orc_volume *v, float bounds[6], float *cell_size, int64_t *kept,
                          int64_t *skipped) {
  orc::TetField f;
  if (v->field_type != 2 || !f.open(v)) return -1;
  for (int a = 0; a < 3; ++a) {
    bounds[a] = f.blo[a];
    bounds[3 + a] = f.bhi[a];
  }
  *cell_size = f.cell_size;
  *kept = int64_t(f.cells.size());
  *skipped = f.skipped;
  return 0;
}

void orc_post_classify(const orc_volume *v, int n, const float *vals, float *out) {
  Volume<float> vol(v);
  for (int i = 0; i < n; ++i) {
    V4<float> c = vol.classify(vals[i]);
    out[4 * i] = c.x;
    out[4 * i + 1] = c.y;
    out[4 * i + 2] = c.z;
    out[4 * i + 3] = c.w;
  }
}

// Conservative majorant grid over voxel space [0,1]^3.  For macrocell m on
// an axis of G cells the samples inside it have texel coordinates
// c in [m*N/G - .5, (m+1)*N/G - .5] and read voxels floor(c) .. floor(c)+1,
// clamped; the cell's value range is the min/max over that voxel block.  The
// TF taps that values in [vmin,vmax] can read are floor(u*n-.5) ..
// floor(u*n-.5)+1; the majorant is the max table alpha over them (the
// reference samples the table at texel centres the same way,
// GridAccel.cpp:69-107).
void orc_build_majorants(const orc_volume *v, int gx, int gy, int gz, float *out) {
  int G[3] = {gx, gy, gz};
  if (v->field_type == 1) {
    orc::build_majorants_nanovdb(v, G, out);
    return;
  }
  if (v->field_type == 2) {
    orc::build_majorants_unstructured(v, G, out);
    return;
  }
  const int *N = v->dims;
  auto range = [&](int a, int m, int &lo, int &hi) {
    double c0 = double(m) * N[a] / G[a] - 0.5;
    double c1 = double(m + 1) * N[a] / G[a] - 0.5;
    lo = std::max(0, int(std::floor(c0)));
    hi = std::min(N[a] - 1, int(std::floor(c1)) + 1);
  };
  int n = v->tf_size;
  for (int mz = 0; mz < gz; ++mz)
    for (int my = 0; my < gy; ++my)
      for (int mx = 0; mx < gx; ++mx) {
        int l[3], h[3];
        range(0, mx, l[0], h[0]);
        range(1, my, l[1], h[1]);
        range(2, mz, l[2], h[2]);
        float vmin = FLT_MAX, vmax = -FLT_MAX;
        for (int z = l[2]; z <= h[2]; ++z)
          for (int y = l[1]; y <= h[1]; ++y)
            for (int x = l[0]; x <= h[0]; ++x) {
              float val = v->voxels[(size_t(z) * N[1] + y) * N[0] + x];
              vmin = std::min(vmin, val);
              vmax = std::max(vmax, val);
            }
        float u0 = (vmin - v->value_lo) / (v->value_hi - v->value_lo);
        float u1 = (vmax - v->value_lo) / (v->value_hi - v->value_lo);
        float c0 = std::min(std::max(u0 * n - 0.5f, -1.f), float(n));
        float c1 = std::min(std::max(u1 * n - 0.5f, -1.f), float(n));
        int lo = std::min(std::max(int(std::floor(c0)), 0), n - 1);
        int hi = std::min(std::max(int(std::floor(c1)) + 1, 0), n - 1);
        float m = 0.f;
        for (int i = lo; i <= hi; ++i) m = std::max(m, v->tf_rgba[4 * i + 3]);
        out[(size_t(mz) * gy + my) * gx + mx] = quantize_majorant(m, majorant_step(v));
      }
}

// Restatement of the reference CPU grid build for comparison:
// StructuredRegularField::buildGrid (:95-127) -> rasterizeBox (GridAccel.h:51-67)
// -> computeMaxOpacities (GridAccel.cpp:69-107).
void orc_build_majorants_reference(const orc_volume *v, int gx, int gy, int gz, float *out) {
  Volume<float> vol(v);
  int G[3] = {gx, gy, gz};
  size_t nmc = size_t(gx) * gy * gz;
  std::vector<float> vmin(nmc, FLT_MAX), vmax(nmc, -FLT_MAX);
  auto project = [&](V3<float> P, int *c) {
    for (int a = 0; a < 3; ++a) {
      float v01 = (P[a] - vol.blo[a]) / (vol.bhi[a] - vol.blo[a]);
      int ci = int(v01 * float(G[a]));
      c[a] = std::min(std::max(ci, 0), G[a] - 1);
    }
  };
  for (int z = 0; z < v->dims[2]; ++z)
    for (int y = 0; y < v->dims[1]; ++y)
      for (int x = 0; x < v->dims[0]; ++x) {
        V3<float> P(v->origin[0] + x * v->spacing[0], v->origin[1] + y * v->spacing[1],
                    v->origin[2] + z * v->spacing[2]);
        float value = vol.sample_regular(vol.point_to_voxel(P));
        V3<float> Q(v->origin[0] + (x + 1) * v->spacing[0], v->origin[1] + (y + 1) * v->spacing[1],
                    v->origin[2] + (z + 1) * v->spacing[2]);
        int lo[3], hi[3];
        project(P, lo);
        project(Q, hi);
        for (int mz = lo[2]; mz <= hi[2]; ++mz)
          for (int my = lo[1]; my <= hi[1]; ++my)
            for (int mx = lo[0]; mx <= hi[0]; ++mx) {
              size_t id = (size_t(mz) * gy + my) * gx + mx;
              vmin[id] = std::min(vmin[id], value);
              vmax[id] = std::max(vmax[id], value);
            }
      }
  int n = v->tf_size;
  for (size_t id = 0; id < nmc; ++id) {
    if (vmax[id] < vmin[id]) { out[id] = 0.f; continue; }
    float lo = (vmin[id] - v->value_lo) / (v->value_hi - v->value_lo);
    float hi = (vmax[id] - v->value_lo) / (v->value_hi - v->value_lo);
    int l = std::min(std::max(int(lo * (n - 1)), 0), n - 1);
    int h = std::min(std::max(int(hi * (n - 1)) + 1, 0), n - 1);
    float m = 0.f;
    for (int i = l; i <= h; ++i) m = std::max(m, v->tf_rgba[4 * i + 3]);
    out[id] = m;
  }
}

int orc_dda_cells(const float ori[3], const float dir[3], float tmin, float tmax,
                  const int dims[3], const float lo[3], const float hi[3], int max_cells,
                  int *cells, float *t0s, float *t1s) {
  Ray<float> r;
  r.ori = from<float>(ori);
  r.dir = from<float>(dir);
  r.tmin = tmin;
  r.tmax = tmax;
  int n = 0;
  dda3(r, dims, from<float>(lo), from<float>(hi), [&](int c, float a, float b) {
    if (n >= max_cells) return false;
    cells[n] = c;
    t0s[n] = a;
    t1s[n] = b;
    ++n;
    return true;
  });
  return n;
}

int orc_delta_track(const orc_volume *v, const float ori[3], const float dir[3], float tmin,
                    float tmax, uint32_t *state, float *t_out, int fp64) {
  Random r;
  r.state = *state;
  int hit;
  if (fp64) {
    Volume<double> vol(v);
    Ray<double> ray{from<double>(ori), from<double>(dir), tmin, tmax};
    VolHit<double> h = delta_track(vol, vol.to_local(ray), r, nullptr);
    hit = h.hit;
    *t_out = float(h.t);
  } else {
    Volume<float> vol(v);
    Ray<float> ray{from<float>(ori), from<float>(dir), tmin, tmax};
    VolHit<float> h = delta_track(vol, vol.to_local(ray), r, nullptr);
    hit = h.hit;
    *t_out = h.t;
  }
  *state = r.state;
  return hit;
}

float orc_ratio_track(const orc_volume *v, const float ori[3], const float dir[3], float tmin,
                      float tmax, uint32_t *state, int fp64) {
  Random r;
  r.state = *state;
  float tr;
  if (fp64) {
    Volume<double> vol(v);
    Ray<double> ray{from<double>(ori), from<double>(dir), tmin, tmax};
    tr = float(delta_track(vol, vol.to_local(ray), r, nullptr, true).Tr);
  } else {
    Volume<float> vol(v);
    Ray<float> ray{from<float>(ori), from<float>(dir), tmin, tmax};
    tr = delta_track(vol, vol.to_local(ray), r, nullptr, true).Tr;
  }
  *state = r.state;
  return tr;
}

void orc_camera_ray(const orc_camera *c, float x, float y, int W, int H, float *ori, float *dir) {
  Camera<float> cam(c);
  Ray<float> r = cam.primary(x, y, float(W), float(H));
  for (int i = 0; i < 3; ++i) {
    ori[i] = r.ori[i];
    dir[i] = r.dir[i];
  }
}

void orc_camera_lens_ray(const orc_camera *c, float x, float y, int W, int H, float u1, float u2,
                         double *ori, double *dir) {
  Camera<double> cam(c);
  Ray<double> r = cam.primary(double(x), double(y), double(W), double(H));
  thin_lens(cam, r, double(u1), double(u2));
  for (int i = 0; i < 3; ++i) {
    ori[i] = r.ori[i];
    dir[i] = r.dir[i];
  }
}

// Frame::fillGBuffer (DeviceCopyableObjects.h:2745-2774) of one pixel
static void fill_gbuffer(const orc::GBuf<float> &g, size_t i, float *depth, float *normal,
                         float *albedo, uint32_t *prim, uint32_t *obj, uint32_t *inst) {
  if (g.depth > depth[i]) return;
  depth[i] = g.depth;
  if (normal) {
    normal[3 * i] = g.Ng.x;
    normal[3 * i + 1] = g.Ng.y;
    normal[3 * i + 2] = g.Ng.z;
  }
  if (albedo) {
    albedo[3 * i] = g.albedo.x;
    albedo[3 * i + 1] = g.albedo.y;
    albedo[3 * i + 2] = g.albedo.z;
  }
  if (prim) prim[i] = g.prim;
  if (obj) obj[i] = g.obj;
  if (inst) inst[i] = g.inst;
}

void orc_pathtrace_gbuffer(const orc_scene *scene, const orc_camera *c, const orc_params *P,
                           int W, int H, int accumID, float *depth, float *normal, float *albedo,
                           uint32_t *prim, uint32_t *obj, uint32_t *inst, int threads) {
  Scene<float> sc(scene);
  Camera<float> cam(c);
  parallel_tiles(W, H, threads, [&](int x, int y) {
    GBuf<float> g;
    (void)pathtrace_pixel(sc, cam, P, x, y, W, H, accumID, nullptr, &g);
    fill_gbuffer(g, size_t(y) * W + x, depth, normal, albedo, prim, obj, inst);
  });
}

// Raycast_impl.cpp:18-20, 135-143, 230-237: the first surface's depth,
// normal, albedo and ids (else depth 1e31, ids unset); a hit of the closest
// volume box then sets the object and instance ids, normal and albedo 0
void orc_raycast_gbuffer(const orc_scene *scene, const orc_camera *c, const orc_params *P,
                         int W, int H, float *depth, float *normal, float *albedo,
                         uint32_t *prim, uint32_t *obj, uint32_t *inst, int threads) {
  Scene<float> sc(scene);
  Camera<float> cam(c);
  parallel_tiles(W, H, threads, [&](int x, int y) {
    GBuf<float> g;
    (void)raycast_pixel(sc, cam, P, x, y, W, H, &g);
    fill_gbuffer(g, size_t(y) * W + x, depth, normal, albedo, prim, obj, inst);
  });
}

// the camera's proj * view, row-major: perspective(fovy, aspect, .001, 1000)
// and the view of eye / dir / up (camera/Perspective.cpp:36-40; visionaray's
// pinhole matrices, parity unpinned: gluPerspective, gluLookAt)
static void camera_view_proj(const orc_camera *c, float M[16]) {
  auto nrm = [](V3<float> a) { return a * (1.f / length(a)); };
  const V3<float> f = nrm(from<float>(c->dir)), up = from<float>(c->up), eye = from<float>(c->pos);
  const V3<float> s = nrm(cross(f, up));
  const V3<float> u = cross(s, f);
  const float V[16] = {s.x, s.y, s.z, -dot(s, eye), u.x, u.y, u.z, -dot(u, eye),
                       -f.x, -f.y, -f.z, dot(f, eye), 0.f, 0.f, 0.f, 1.f};
  const float zn = .001f, zf = 1000.f;
  const float g = 1.f / std::tan(c->fovy * 0.5f);
  const float Pm[16] = {g / c->aspect, 0.f, 0.f, 0.f, 0.f, g, 0.f, 0.f,
                        0.f, 0.f, (zf + zn) / (zn - zf), (2.f * zf * zn) / (zn - zf),
                        0.f, 0.f, -1.f, 0.f};
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k) {
      float acc = 0.f;
      for (int j = 0; j < 4; ++j) acc += Pm[4 * r + j] * V[4 * j + k];
      M[4 * r + k] = acc;
    }
}

// One TAA frame (Pathtrace_impl.cpp:636-637, Frame.cpp:339-381,
// DeviceCopyableObjects.h:2657-2712), FP32: every pixel's mean sample, the
// closest sample's albedo and last-vertex motion vector (0 for a pixel
// without hits), then the resolve against the history `prev_rgba` /
// `prev_albedo` (read and replaced by this frame's, as the reference copies
// them) into `out_rgba`.  prev_camera NULL: no motion.
void orc_pathtrace_taa(const orc_scene *scene, const orc_camera *c, const orc_camera *prev_camera,
                       const orc_params *P, int W, int H, int accumID, float alpha,
                       float *prev_rgba, float *prev_albedo, float *out_rgba, float *motion,
                       int threads) {
  Scene<float> sc(scene);
  Camera<float> cam(c);
  float vp_curr[16], vp_prev[16];
  camera_view_proj(c, vp_curr);
  if (prev_camera) camera_view_proj(prev_camera, vp_prev);
  else std::copy(vp_curr, vp_curr + 16, vp_prev);
  const size_t n = size_t(W) * H;
  std::vector<float> curr(4 * n), calb(3 * n), mv(2 * n);
  parallel_tiles(W, H, threads, [&](int x, int y) {
    GBuf<float> g;
    g.vp_prev = vp_prev;
    g.vp_curr = vp_curr;
    g.W = W;
    g.H = H;
    const V4<float> s = pathtrace_pixel(sc, cam, P, x, y, W, H, accumID, nullptr, &g);
    const size_t i = size_t(y) * W + x;
    curr[4 * i] = s.x;
    curr[4 * i + 1] = s.y;
    curr[4 * i + 2] = s.z;
    curr[4 * i + 3] = s.w;
    const bool hit = g.depth < 1e31f;
    mv[2 * i] = hit ? g.mvx : 0.f;
    mv[2 * i + 1] = hit ? g.mvy : 0.f;
    for (int k = 0; k < 3; ++k) calb[3 * i + k] = hit ? g.albedo[k] : 0.f;
  });
  parallel_tiles(W, H, threads, [&](int x, int y) {
    const size_t i = size_t(y) * W + x;
    const float mx = mv[2 * i], my = mv[2 * i + 1];
    int px = int(float(x) + mx), py = int(float(y) + my);
    px = std::min(std::max(px, 0), W - 1);
    py = std::min(std::max(py, 0), H - 1);
    const size_t pi = size_t(py) * W + px;
    float a = alpha;
    if (!(std::fabs(prev_albedo[3 * pi] - calb[3 * i]) < 1e-2f &&
          std::fabs(prev_albedo[3 * pi + 1] - calb[3 * i + 1]) < 1e-2f &&
          std::fabs(prev_albedo[3 * pi + 2] - calb[3 * i + 2]) < 1e-2f))
      a = 1.f;
    // tex2D(history, ((x + mx + .5) / W, (y + my + .5) / H)): bilinear, clamp
    // (parity unpinned: visionaray's tex2D)
    const float cx = std::fmin(std::fmax(float(x) + mx, -1.f), float(W));
    const float cy = std::fmin(std::fmax(float(y) + my, -1.f), float(H));
    const float fx = std::floor(cx), fy = std::floor(cy);
    const float wx = cx - fx, wy = cy - fy;
    const int x0 = std::min(std::max(int(fx), 0), W - 1), x1 = std::min(std::max(int(fx) + 1, 0), W - 1);
    const int y0 = std::min(std::max(int(fy), 0), H - 1), y1 = std::min(std::max(int(fy) + 1, 0), H - 1);
    auto lp = [](float p, float q, float t) { return p + t * (q - p); };
    const float b = 1.f - a;
    for (int k = 0; k < 4; ++k) {
      const float h = lp(lp(prev_rgba[4 * (size_t(y0) * W + x0) + k], prev_rgba[4 * (size_t(y0) * W + x1) + k], wx),
                         lp(prev_rgba[4 * (size_t(y1) * W + x0) + k], prev_rgba[4 * (size_t(y1) * W + x1) + k], wx),
                         wy);
      out_rgba[4 * i + k] = b * h + a * curr[4 * i + k];
    }
  });
  std::copy(out_rgba, out_rgba + 4 * n, prev_rgba);
  std::copy(calb.begin(), calb.end(), prev_albedo);
  if (motion) std::copy(mv.begin(), mv.end(), motion);
}

void orc_raycast(const orc_scene *scene, const orc_camera *c, const orc_params *P, int W,
                 int H, float *out, int threads) {
  Scene<float> sc(scene);
  Camera<float> cam(c);
  parallel_tiles(W, H, threads, [&](int x, int y) {
    V4<float> s = raycast_pixel(sc, cam, P, x, y, W, H);
    size_t i = size_t(y) * W + x;
    out[4 * i] = s.x;
    out[4 * i + 1] = s.y;
    out[4 * i + 2] = s.z;
    out[4 * i + 3] = s.w;
  });
}

}  // extern "C"

namespace orc {
template <class R>
static void pt_frame(const orc_scene *scene, const orc_camera *c, const orc_params *P, int W,
                     int H, int accumID, int stride, int y0, int y1, float *out, int threads,
                     uint64_t *lookups) {
  Scene<R> sc(scene);
  Camera<R> cam(c);
  std::atomic<uint64_t> total(0);
  if (y1 <= 0 || y1 > H) y1 = H;
  int rows = y1 - y0;
  parallel_tiles(W, rows, threads, [&](int x, int yy) {
    int y = yy + y0;
    if (stride > 1 && ((x % stride) != 0 || (y % stride) != 0)) return;
    Counter cnt;
    V4<R> s = pathtrace_pixel(sc, cam, P, x, y, W, H, accumID, &cnt);
    size_t i = stride > 1 ? size_t(y / stride) * ((W + stride - 1) / stride) + x / stride
                          : size_t(y) * W + x;
    out[4 * i] = float(s.x);
    out[4 * i + 1] = float(s.y);
    out[4 * i + 2] = float(s.z);
    out[4 * i + 3] = float(s.w);
    total.fetch_add(cnt.lookups, std::memory_order_relaxed);
  });
  if (lookups) *lookups = total.load();
}

// Running mean of Frame::accumSample (DeviceCopyableObjects.h:2713-2716):
// FP32 mode reproduces the float recurrence acc = (1-a)*acc + a*c with
// a = 1/(accumID+1); FP64 mode accumulates the exact double mean.
template <class R>
static void pt_accum(const orc_scene *scene, const orc_camera *c, const orc_params *P, int W,
                     int H, int first, int frames, int stride, double *out, int threads) {
  Scene<R> sc(scene);
  Camera<R> cam(c);
  parallel_tiles(W, H, threads, [&](int x, int y) {
    if (stride > 1 && ((x % stride) != 0 || (y % stride) != 0)) return;
    size_t i = stride > 1 ? size_t(y / stride) * ((W + stride - 1) / stride) + x / stride
                          : size_t(y) * W + x;
    R acc[4] = {0, 0, 0, 0};
    for (int f = 0; f < frames; ++f) {
      int accumID = first + f;
      V4<R> s = pathtrace_pixel(sc, cam, P, x, y, W, H, accumID, nullptr);
      R a = R(1) / R(f + 1);
      R v[4] = {s.x, s.y, s.z, s.w};
      for (int k = 0; k < 4; ++k) acc[k] = (R(1) - a) * acc[k] + a * v[k];
    }
    for (int k = 0; k < 4; ++k) out[4 * i + k] = double(acc[k]);
  });
}
}  // namespace orc

extern "C" {

void orc_pathtrace_frame(const orc_scene *scene, const orc_camera *c, const orc_params *P,
                         int W, int H, int accumID, int frameID, int fp64, int stride, int y0,
                         int y1, float *out, int threads, uint64_t *lookups) {
  (void)frameID;
  if (fp64)
    pt_frame<double>(scene, c, P, W, H, accumID, stride, y0, y1, out, threads, lookups);
  else
    pt_frame<float>(scene, c, P, W, H, accumID, stride, y0, y1, out, threads, lookups);
}

void orc_pathtrace_accumulate(const orc_scene *scene, const orc_camera *c, const orc_params *P,
                              int W, int H, int first, int frames, int fp64, int stride,
                              double *out, int threads) {
  if (fp64)
    pt_accum<double>(scene, c, P, W, H, first, frames, stride, out, threads);
  else
    pt_accum<float>(scene, c, P, W, H, first, frames, stride, out, threads);
}

// frame/common.h:31-47
static uint32_t cvt1(float f) { return uint32_t(255.f * std::min(std::max(f, 0.f), 1.f)); }
uint32_t orc_cvt_uint32(const float v[4]) {
  return (cvt1(v[0]) << 0) | (cvt1(v[1]) << 8) | (cvt1(v[2]) << 16) | (cvt1(v[3]) << 24);
}
// linear_to_srgb (parity unpinned: IEC 61966-2-1 piecewise curve)
static float l2s(float x) {
  if (x <= 0.0031308f) return 12.92f * x;
  return 1.055f * std::pow(x, 1.f / 2.4f) - 0.055f;
}
uint32_t orc_cvt_uint32_srgb(const float v[4]) {
  float c[4] = {l2s(v[0]), l2s(v[1]), l2s(v[2]), v[3]};
  return orc_cvt_uint32(c);
}

float orc_pow_ieee(float x, float y) { return orc::pow_ieee(x, y); }

}  // extern "C"

extern "C" {

int orc_instance_transforms(const float xf[12], float A[9], float tinv[3], float N[9]) {
  for (int i = 0; i < 12; ++i)
    if (!std::isfinite(xf[i])) return -1;
  double L[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) L[3 * r + c] = double(xf[4 * r + c]);
  // the adjugate over the determinant, cofactor by cofactor
  const double c00 = L[4] * L[8] - L[5] * L[7], c01 = L[3] * L[8] - L[5] * L[6];
  const double c02 = L[3] * L[7] - L[4] * L[6];
  const double det = L[0] * c00 - L[1] * c01 + L[2] * c02;
  if (!(det != 0.0) || !std::isfinite(det)) return -1;
  const double inv[9] = {c00 / det, (L[2] * L[7] - L[1] * L[8]) / det,
                         (L[1] * L[5] - L[2] * L[4]) / det, (L[5] * L[6] - L[3] * L[8]) / det,
                         (L[0] * L[8] - L[2] * L[6]) / det, (L[2] * L[3] - L[0] * L[5]) / det,
                         c02 / det, (L[1] * L[6] - L[0] * L[7]) / det,
                         (L[0] * L[4] - L[1] * L[3]) / det};
  for (int i = 0; i < 9; ++i) {
    A[i] = float(inv[i]);
    if (!std::isfinite(A[i])) return -1;
  }
  for (int k = 0; k < 3; ++k) tinv[k] = -xf[4 * k + 3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) N[3 * r + c] = A[3 * c + r];
  return 0;
}

void orc_scene_intersect_surfaces(const orc_scene *scene, int n, const float *rays, int any,
                                  float *t_out, int *prim_out, int *surf_out) {
  const orc::Scene<float> sc(scene);
  for (int i = 0; i < n; ++i) {
    const float *r = rays + 8 * size_t(i);
    orc::Ray<float> ray;
    ray.ori = orc::V3<float>(r[0], r[1], r[2]);
    ray.dir = orc::V3<float>(r[3], r[4], r[5]);
    ray.tmin = r[6];
    ray.tmax = r[7];
    float t, u, v;
    int id = -1, prim = -1;
    const bool hit = sc.surfaces_hit(ray, any != 0, t, id, prim, u, v);
    t_out[i] = hit ? t : std::numeric_limits<float>::infinity();
    prim_out[i] = hit ? prim : -1;
    surf_out[i] = hit ? id : -1;
  }
}

void orc_mesh_intersect(const orc_trimesh *m, int n, const float *rays, int any, int brute,
                        float *t_out, int *prim_out) {
  orc::Mesh<float> mesh(m);
  for (int i = 0; i < n; ++i) {
    const float *r = rays + 8 * size_t(i);
    orc::Ray<float> ray;
    ray.ori = orc::V3<float>(r[0], r[1], r[2]);
    ray.dir = orc::V3<float>(r[3], r[4], r[5]);
    ray.tmin = r[6];
    ray.tmax = r[7];
    float t = std::numeric_limits<float>::infinity(), u, v;
    int prim = -1;
    bool hit = false;
    if (brute) {
      float best = ray.tmax;
      for (size_t k = 0; k < mesh.tris.size(); ++k) {
        const orc::TriHit<float> h = orc::intersect_tri(ray, mesh.tris[k]);
        if (!h.hit || h.t < ray.tmin || h.t > best) continue;
        if (hit && h.t == best && int(k) > prim) continue;
        hit = true;
        best = h.t;
        prim = int(k);
        if (any) break;
      }
      if (hit) t = best;
    } else {
      float tb;
      hit = mesh.intersect(ray, any != 0, tb, prim, u, v);
      if (hit) t = tb;
    }
    t_out[i] = hit ? t : std::numeric_limits<float>::infinity();
    prim_out[i] = hit ? prim : -1;
  }
}

void orc_hg_sample(float g, const float d[3], int n, uint32_t seed, float *out) {
  const V3<float> dir = normalize(V3<float>(d[0], d[1], d[2]));
  for (int i = 0; i < n; ++i) {
    Random r(seed, uint32_t(i));
    const float u1 = r(), u2 = r();
    const V3<float> w = hg_sample(g, dir, u1, u2);
    out[3 * i] = w.x;
    out[3 * i + 1] = w.y;
    out[3 * i + 2] = w.z;
  }
}

double orc_hg_4pi(double g, double cos_theta) { return hg_4pi(g, cos_theta); }

void orc_uniform_sample_sphere(int n, uint32_t seed, float *out) {
  for (int i = 0; i < n; ++i) {
    Random r(seed, uint32_t(i));
    const float u1 = r(), u2 = r();
    const V3<float> w = uniform_sample_sphere(u1, u2);
    out[3 * i] = w.x;
    out[3 * i + 1] = w.y;
    out[3 * i + 2] = w.z;
  }
}

void orc_set_roulette(int enabled) { g_roulette = enabled != 0; }

int orc_hdri_cdf(const orc_light *L, float *rows, float *cols) {
  orc::HdriLight H(*L);
  if (!H.ok) return -1;
  std::copy(H.rows.begin(), H.rows.end(), rows);
  std::copy(H.cols.begin(), H.cols.end(), cols);
  return 0;
}

void orc_hdri_sample(const orc_light *L, int n, uint32_t seed, float *dirs, float *pdfs,
                     float *rgb) {
  orc::HdriLight H(*L);
  for (int i = 0; i < n; ++i) {
    orc::Random rnd(seed, uint32_t(i));
    orc::V3<float> d;
    float pdf;
    H.sample(rnd, d, pdf);
    const orc::V3<float> I = H.radiance(d);
    for (int k = 0; k < 3; ++k) {
      dirs[3 * i + k] = d[k];
      rgb[3 * i + k] = I[k];
    }
    pdfs[i] = pdf;
  }
}

void orc_hdri_eval(const orc_light *L, int n, const float *dirs, float *rgb, float *pdfs,
                   float *uv) {
  orc::HdriLight H(*L);
  for (int i = 0; i < n; ++i) {
    const orc::V3<float> d(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
    const orc::V3<float> I = H.radiance(d);
    for (int k = 0; k < 3; ++k) rgb[3 * i + k] = I[k];
    pdfs[i] = H.pdf(d);
    orc::HdriLight::to_uv(orc::HdriLight::mul(H.L, d), uv[2 * i], uv[2 * i + 1]);
  }
}

void orc_quad_light_sample(const orc_light *L, const float ref[3], int n, uint32_t seed,
                           float *pos_out, float *pdf_out) {
  orc::QuadLight<float> Q(*L);
  orc::HdriLight H(*L);
  const orc::V3<float> p(ref[0], ref[1], ref[2]);
  for (int i = 0; i < n; ++i) {
    orc::Random rnd(seed, uint32_t(i));
    const orc::LightSample<float> ls = orc::sample_light<float>(*L, Q, H, p, rnd);
    const orc::V3<float> pos = p + ls.dir;
    pos_out[3 * i] = pos.x;
    pos_out[3 * i + 1] = pos.y;
    pos_out[3 * i + 2] = pos.z;
    pdf_out[i] = ls.pdf;
  }
}

}  // extern "C"
