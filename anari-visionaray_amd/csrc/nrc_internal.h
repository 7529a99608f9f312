// The radiance-cache object shared by the cache ABI (nrc_api.cpp) and the
// path tracer entry point (vrt_api.cpp).
#pragma once

#include "abi_internal.h"
#include "kernel_common.h"
#include "nrc_kernels.h"

struct vrt_nrc_s {
  vrt_context ctx = nullptr;
  vrt_nrc_config cfg{};
  vrt::NrcDevice d{};
  int max_workgroups = 0;
  // per-frame buffers of vrt_pathtrace_nrc, grown on demand
  size_t pixels = 0;
  int t_cap = 0;
  float4 *sample = nullptr;
  int *query_of_pixel = nullptr;
  float *q_raw = nullptr, *q_T = nullptr, *q_out = nullptr;
  float *t_raw = nullptr, *t_target = nullptr, *t_ratio = nullptr;
  int *t_query = nullptr;
  uint32_t *ctr = nullptr;  // see vrt::KNrc::ctr
};

namespace vrt {

// make sure the frame buffers hold `pixels` queries and `t_cap` training
// samples (allocation happens only when they grow)
int nrc_frame_buffers(vrt_nrc n, size_t pixels, int t_cap);
// `steps` Adam steps over this frame's training samples (count on the GPU)
int nrc_train_frame(vrt_nrc n, int steps);

}  // namespace vrt
