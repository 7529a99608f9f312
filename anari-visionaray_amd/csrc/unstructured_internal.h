// The tetrahedral field behind vrt_unstructured (unstructured_field.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "abi_internal.h"

struct vrt_unstructured_s {
  int device = 0;
  // device: point-location grid (cell list starts, dims product + 1 entries;
  // 64-byte tetrahedron copies {x, y, z, value} x 4 in list order) and, per
  // tetrahedron, its bounding box and value range ({min xyz, min value},
  // {max xyz, max value}) for the majorant grid
  uint32_t *start = nullptr;
  float4 *recs = nullptr;
  uint32_t *rec_cell = nullptr;  // per list entry: the ANARI cell index (primitive id)
  float4 *tets = nullptr;
  float lo[3] = {0, 0, 0}, inv_cell[3] = {1, 1, 1};
  int dims[3] = {1, 1, 1};
  float blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
  float cell_size = 1.f;
  uint64_t num_tets = 0, skipped = 0, entries = 0, device_bytes = 0;
  uint32_t max_list = 0;
  double mean_list = 0, expected_tests = 0, build_ms = 0;
};
