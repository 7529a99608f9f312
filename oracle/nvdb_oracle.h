// ORACLE -- test infrastructure only (see vrt_oracle.h).
//
// NanoVDB float grids (file format 32.7, external/nanovdb/NanoVDB.h): value
// lookup by walking the tree of a raw grid buffer (root tiles -> upper
// internal node -> lower internal node -> leaf), and the samplers the
// reference's sampleField uses (DeviceCopyableObjects.h:554-569 with
// math/SampleFromVoxels.h:174-290, 324-430).  Restated from the format's
// definition; pinned against the vendored NanoVDB in oracle/_ref.
#pragma once
#include <cstdint>

namespace orc {

struct NanoVDBGrid {
  const uint8_t *buf = nullptr;
  float inv_f[9], vec_f[3];  // Map::mInvMatF, mVecF (world -> index, float)
  double inv_d[9], vec_d[3]; // Map::mInvMatD, mVecD
  double wlo[3], whi[3];     // GridData::mWorldBBox
  float background = 0.f;
  const uint8_t *root = nullptr;
  uint32_t tiles = 0;

  // false when the buffer is not a float grid of a compatible version
  bool open(const void *buffer);
  float get(int i, int j, int k) const;  // ReadAccessor::getValue
  // createSampler<order>(acc)(worldToIndexF(P)); order 0 nearest, 1 linear
  float sample(const float P[3], int order) const;
  double sample(const double P[3], int order) const;
};

}  // namespace orc
