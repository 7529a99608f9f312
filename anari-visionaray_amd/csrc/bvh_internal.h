// The triangle BVH behind vrt_bvh and the instance BVH behind vrt_tls
// (bvh_build.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernel_common.h"

struct vrt_bvh_s {
  int device = 0;
  // device: 4-wide nodes (8 float4 each: child boxes lo/hi x, y, z as four
  // lanes, child entries, leaf counts) and triangles (3 float4: v1 + prim id
  // bits, e1, e2) in leaf order
  float4 *nodes = nullptr;
  float4 *tris = nullptr;
  uint32_t num_nodes = 0, num_tris = 0, depth = 0;
  uint32_t stack_bound = 0;  // most traversal-stack entries (vrt_bvh_info)
  float lo[3], hi[3];
  double build_ms = 0.0;
};

struct vrt_tls_s {
  int device = 0;
  float4 *nodes = nullptr;     // 4-wide nodes, leaves index `inst`
  vrt::KInst *inst = nullptr;  // instance records in leaf order
  uint32_t num_nodes = 0, num_inst = 0, stack_bound = 0;
  float lo[3], hi[3];
  double build_ms = 0.0;
};
