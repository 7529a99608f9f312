// Test infrastructure only: golden NanoVDB behaviour from the reference's
// vendored NanoVDB 32.7 (external/nanovdb, header-only), compiled by
// oracle/build.py into oracle/_ref/ when /root/reference is present.  It
// validates the grids our generator writes (tools::checkGrid), answers
// getValue / createSampler<0|1> queries as NanoVDBField's sampleField does
// (DeviceCopyableObjects.h:554-569), and builds grids with tiles at every
// level through NanoVDB's own builder, which our writer never emits.
#include <nanovdb/NanoVDB.h>
#include <nanovdb/math/SampleFromVoxels.h>
#include <nanovdb/tools/CreateNanoGrid.h>
#include <nanovdb/tools/GridBuilder.h>
#include <nanovdb/tools/GridValidator.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>

using GridF = nanovdb::NanoGrid<float>;

extern "C" {

// 0 when the buffer holds a valid float grid (full structural check)
int ref_nvdb_check(const void *buf, char *err, int errlen) {
  const auto *g = reinterpret_cast<const GridF *>(buf);
  char msg[512];
  nanovdb::tools::checkGrid(g, msg, nanovdb::CheckMode::Full);
  if (err && errlen > 0) {
    std::strncpy(err, msg, size_t(errlen) - 1);
    err[errlen - 1] = '\0';
  }
  return msg[0] ? 1 : 0;
}

void ref_nvdb_get(const void *buf, int n, const int32_t *ijk, float *out) {
  const auto *g = reinterpret_cast<const GridF *>(buf);
  for (int i = 0; i < n; ++i) {
    auto acc = g->getAccessor();
    out[i] = acc.getValue(nanovdb::Coord(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]));
  }
}

// sampleField of a nanovdb field: a fresh accessor and sampler per point,
// position mapped with worldToIndexF
void ref_nvdb_sample(const void *buf, int n, const float *xyz, int order, float *out) {
  const auto *g = reinterpret_cast<const GridF *>(buf);
  for (int i = 0; i < n; ++i) {
    auto acc = g->getAccessor();
    nanovdb::math::Vec3<float> p(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    if (order == 0) {
      auto smp = nanovdb::math::createSampler<0>(acc);
      out[i] = smp(g->worldToIndexF(p));
    } else {
      auto smp = nanovdb::math::createSampler<1>(acc);
      out[i] = smp(g->worldToIndexF(p));
    }
  }
}

void ref_nvdb_world_bbox(const void *buf, double *out6) {
  const auto *g = reinterpret_cast<const GridF *>(buf);
  const auto bb = g->worldBBox();
  for (int a = 0; a < 3; ++a) {
    out6[a] = bb.min()[a];
    out6[3 + a] = bb.max()[a];
  }
}

// A grid with an active root tile, upper and lower internal-node tiles,
// leaves next to them (including negative coordinates) and a scaled,
// translated map.  Returns a malloc'd buffer (free with ref_nvdb_free).
int ref_nvdb_make_tiled(uint64_t seed, uint8_t **out, uint64_t *bytes) {
  nanovdb::tools::build::Grid<float> grid(0.25f, "tiled", nanovdb::GridClass::FogVolume);
  grid.setTransform(0.5, nanovdb::Vec3d(-3.0, 1.0, 2.0));
  auto &root = grid.tree().root();
  root.template addTile<3>(nanovdb::Coord(4096, 0, 0), 0.7f, true);     // 4096^3 root tile
  root.template addTile<2>(nanovdb::Coord(128, 0, 0), 0.4f, true);      // 128^3 upper tile
  root.template addTile<1>(nanovdb::Coord(-8, -8, -8), 0.9f, false);    // 8^3 lower tile
  auto acc = grid.getAccessor();
  uint64_t s = seed ? seed : 1;
  auto next = [&]() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return float(s >> 40) / float(1 << 24);
  };
  for (int k = -12; k < 20; ++k)
    for (int j = -12; j < 20; ++j)
      for (int i = -12; i < 20; ++i) {
        if ((i >= -8 && i < 0) && (j >= -8 && j < 0) && (k >= -8 && k < 0)) continue;
        if (next() < 0.35f) acc.setValue(nanovdb::Coord(i, j, k), next());
      }
  for (int k = 0; k < 6; ++k)
    for (int j = 0; j < 6; ++j)
      for (int i = 120; i < 136; ++i)
        if (i < 128 && next() < 0.5f) acc.setValue(nanovdb::Coord(i, j, k), 2.f * next());
  for (int k = 0; k < 4; ++k)
    for (int i = 4090; i < 4096; ++i) acc.setValue(nanovdb::Coord(i, 3, k), next());
  auto handle = nanovdb::tools::createNanoGrid(grid);
  *bytes = handle.size();
  *out = static_cast<uint8_t *>(std::malloc(handle.size()));
  if (!*out) return -1;
  std::memcpy(*out, handle.data(), handle.size());
  return 0;
}

void ref_nvdb_free(void *p) { std::free(p); }

}  // extern "C"
