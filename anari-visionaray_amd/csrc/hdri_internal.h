// The HDRI environment map behind vrt_hdri (hdri_light.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "kernel_common.h"

struct vrt_hdri_s {
  int device = 0;
  int w = 0, h = 0;
  float4 *tex = nullptr;  // RGBA radiance (makeRGBA, HDRI.cpp:72-89)
  float *rows = nullptr;  // CDF over the row totals
  float *cols = nullptr;  // per-row CDFs
};

namespace vrt {
// the light's frame (HDRI::finalize, HDRI.cpp:150-155): col1 = -up,
// col0 = col1 x direction, col2 = col0 x col1, each normalised; to_local its
// inverse.  False if the frame is degenerate.
bool hdri_frame(const float up[3], const float direction[3], float to_world[9],
                float to_local[9]);
}  // namespace vrt
