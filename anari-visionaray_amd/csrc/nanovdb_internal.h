// The converted NanoVDB grid behind vrt_nanovdb (nanovdb_field.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "abi_internal.h"

struct vrt_nanovdb_s {
  int device = 0;
  hipStream_t stream = nullptr;  // the creating context's stream (not used after destroy)
  // device: top table over 128^3 regions (block index >= 0, or -1 - constant),
  // blocks of 16^3 slot entries (leaf index >= 0, or -1 - constant), leaf
  // values (512 floats per leaf), constants (background first), per-slot
  // value ranges of the blocks (float2)
  int32_t *top = nullptr;
  int32_t *blocks = nullptr;
  float *leaves = nullptr;
  float *consts = nullptr;
  float *slot_ranges = nullptr;
  int lo[3] = {0, 0, 0}, n[3] = {0, 0, 0};  // top-table origin (in 128-voxel units) / dims
  uint64_t block_count = 0;
  float inv_f[9], vec_f[3];                  // Map::mInvMatF, mVecF
  double inv_d[9], vec_d[3];                 // Map::mInvMatD, mVecD
  float blo[3], bhi[3];                      // world bbox as float
  float background = 0.f;
  uint64_t leaf_count = 0, const_count = 0, active = 0, device_bytes = 0;
};

namespace vrt {
// voxel-index range [lo, hi] (inclusive, per axis) that trilinear samples at
// any point of macrocell `cell` of a grid_dims grid over [glo, ghi] can read
// (8 corners through the double inverse map, floor, one voxel of margin
// below, two above)
void nanovdb_cell_voxels(const vrt_nanovdb_s &g, const float glo[3], const float ghi[3],
                         const int dims[3], int cx, int cy, int cz, int lo[3], int hi[3]);
}  // namespace vrt
