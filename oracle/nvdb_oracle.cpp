// ORACLE -- test infrastructure only.  NanoVDB grid walk and samplers; see
// nvdb_oracle.h.  Byte offsets are those of the 32.7 records: GridData
// (672 B: map at 296, world bbox at 560, grid type at 636), TreeData (64 B:
// node offsets, root last), RootData (64 B: bbox, tile count at 24,
// background at 28) followed by 32-B tiles (key, child offset from the root,
// state, value), upper / lower internal nodes (child mask after the value
// mask, 8-B table entries holding a float tile value or a child offset from
// the node), leaves (512 floats at 96, z-fastest).
#include "nvdb_oracle.h"

#include <cmath>
#include <cstring>

namespace orc {

namespace {
template <class T> T rd(const uint8_t *p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}
}  // namespace

bool NanoVDBGrid::open(const void *buffer) {
  buf = static_cast<const uint8_t *>(buffer);
  const uint64_t magic = rd<uint64_t>(buf);
  if (magic != 0x304244566f6e614eull && magic != 0x314244566f6e614eull) return false;
  if ((rd<uint32_t>(buf + 16) >> 21) != 32u) return false;  // major version
  if (rd<uint32_t>(buf + 636) != 1u) return false;           // GridType::Float
  for (int i = 0; i < 9; ++i) inv_f[i] = rd<float>(buf + 296 + 36 + 4 * i);
  for (int i = 0; i < 3; ++i) vec_f[i] = rd<float>(buf + 296 + 72 + 4 * i);
  for (int i = 0; i < 9; ++i) inv_d[i] = rd<double>(buf + 296 + 160 + 8 * i);
  for (int i = 0; i < 3; ++i) vec_d[i] = rd<double>(buf + 296 + 232 + 8 * i);
  for (int i = 0; i < 3; ++i) wlo[i] = rd<double>(buf + 560 + 8 * i);
  for (int i = 0; i < 3; ++i) whi[i] = rd<double>(buf + 584 + 8 * i);
  const uint8_t *tree = buf + 672;
  root = tree + rd<int64_t>(tree + 24);
  tiles = rd<uint32_t>(root + 24);
  background = rd<float>(root + 28);
  return true;
}

float NanoVDBGrid::get(int i, int j, int k) const {
  // root: linear search of the tile whose key matches (RootData::probeTile)
  const uint64_t key = uint64_t(uint32_t(k) >> 12) | (uint64_t(uint32_t(j) >> 12) << 21) |
                       (uint64_t(uint32_t(i) >> 12) << 42);
  const uint8_t *tile = nullptr;
  for (uint32_t t = 0; t < tiles; ++t) {
    const uint8_t *p = root + 64 + 32 * size_t(t);
    if (rd<uint64_t>(p) == key) {
      tile = p;
      break;
    }
  }
  if (!tile) return background;
  const int64_t child = rd<int64_t>(tile + 8);
  if (child == 0) return rd<float>(tile + 20);
  // upper internal node, 32^3 children of 128^3 voxels
  const uint8_t *up = root + child;
  const uint32_t nu = uint32_t(((i & 4095) >> 7) << 10 | ((j & 4095) >> 7) << 5 | ((k & 4095) >> 7));
  if (!((up[32 + 4096 + (nu >> 3)] >> (nu & 7)) & 1u)) return rd<float>(up + 8256 + 8 * size_t(nu));
  // lower internal node, 16^3 children of 8^3 voxels
  const uint8_t *lo = up + rd<int64_t>(up + 8256 + 8 * size_t(nu));
  const uint32_t nl = uint32_t(((i & 127) >> 3) << 8 | ((j & 127) >> 3) << 4 | ((k & 127) >> 3));
  if (!((lo[32 + 512 + (nl >> 3)] >> (nl & 7)) & 1u)) return rd<float>(lo + 1088 + 8 * size_t(nl));
  const uint8_t *leaf = lo + rd<int64_t>(lo + 1088 + 8 * size_t(nl));
  return rd<float>(leaf + 96 + 4 * size_t(((i & 7) << 6) | ((j & 7) << 3) | (k & 7)));
}

// TrilinearSampler::sample: lerp over z, then y, then x, lerp(a, b, w) =
// a + w (b - a); Floor leaves the fractional part in the position
float NanoVDBGrid::sample(const float P[3], int order) const {
  // Map::applyInverseMapF: fmaf chain over (P - vec)
  const float x = P[0] - vec_f[0], y = P[1] - vec_f[1], z = P[2] - vec_f[2];
  float idx[3];
  for (int a = 0; a < 3; ++a)
    idx[a] = std::fma(x, inv_f[3 * a], std::fma(y, inv_f[3 * a + 1], z * inv_f[3 * a + 2]));
  if (order == 0)
    return get(int(std::rint(idx[0])), int(std::rint(idx[1])), int(std::rint(idx[2])));
  int c[3];
  float w[3];
  for (int a = 0; a < 3; ++a) {
    const float f = std::floor(idx[a]);
    w[a] = idx[a] - f;
    c[a] = int(f);
  }
  auto lerp = [](float a, float b, float t) { return a + t * (b - a); };
  const int i = c[0], j = c[1], k = c[2];
  const float v00 = lerp(get(i, j, k), get(i, j, k + 1), w[2]);
  const float v01 = lerp(get(i, j + 1, k), get(i, j + 1, k + 1), w[2]);
  const float v10 = lerp(get(i + 1, j, k), get(i + 1, j, k + 1), w[2]);
  const float v11 = lerp(get(i + 1, j + 1, k), get(i + 1, j + 1, k + 1), w[2]);
  return lerp(lerp(v00, v01, w[1]), lerp(v10, v11, w[1]), w[0]);
}

// FP64 restatement for the converged reference: the double map, double
// interpolation weights, float voxel values
double NanoVDBGrid::sample(const double P[3], int order) const {
  const double x = P[0] - vec_d[0], y = P[1] - vec_d[1], z = P[2] - vec_d[2];
  double idx[3];
  for (int a = 0; a < 3; ++a)
    idx[a] = std::fma(x, inv_d[3 * a], std::fma(y, inv_d[3 * a + 1], z * inv_d[3 * a + 2]));
  if (order == 0)
    return get(int(std::floor(idx[0] + 0.5)), int(std::floor(idx[1] + 0.5)),
               int(std::floor(idx[2] + 0.5)));
  int c[3];
  double w[3];
  for (int a = 0; a < 3; ++a) {
    const double f = std::floor(idx[a]);
    w[a] = idx[a] - f;
    c[a] = int(f);
  }
  auto lerp = [](double a, double b, double t) { return a + t * (b - a); };
  const int i = c[0], j = c[1], k = c[2];
  const double v00 = lerp(get(i, j, k), get(i, j, k + 1), w[2]);
  const double v01 = lerp(get(i, j + 1, k), get(i, j + 1, k + 1), w[2]);
  const double v10 = lerp(get(i + 1, j, k), get(i + 1, j, k + 1), w[2]);
  const double v11 = lerp(get(i + 1, j + 1, k), get(i + 1, j + 1, k + 1), w[2]);
  return lerp(lerp(v00, v01, w[1]), lerp(v10, v11, w[1]), w[0]);
}

}  // namespace orc

extern "C" {

int orc_nvdb_get(const void *buf, int n, const int32_t *ijk, float *out) {
  orc::NanoVDBGrid g;
  if (!g.open(buf)) return -1;
  for (int t = 0; t < n; ++t) out[t] = g.get(ijk[3 * t], ijk[3 * t + 1], ijk[3 * t + 2]);
  return 0;
}

int orc_nvdb_sample(const void *buf, int n, const float *xyz, int order, float *out) {
  orc::NanoVDBGrid g;
  if (!g.open(buf)) return -1;
  for (int t = 0; t < n; ++t) out[t] = g.sample(xyz + 3 * t, order);
  return 0;
}

int orc_nvdb_world_bbox(const void *buf, double *out6) {
  orc::NanoVDBGrid g;
  if (!g.open(buf)) return -1;
  for (int a = 0; a < 3; ++a) {
    out6[a] = g.wlo[a];
    out6[3 + a] = g.whi[a];
  }
  return 0;
}

}  // extern "C"
