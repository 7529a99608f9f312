// Internals shared by the C ABI translation units: the context object and
// the error reporting helpers behind vrt_last_error().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vrt.h"

struct vrt_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  int num_cus = 0;
  // work-queue counters of the persistent path tracer: two sets of 8, used
  // alternately; each launch clears the set the next launch uses
  uint32_t *queues = nullptr;
  int queue_parity = 0;
  // device address of the library's file-scope queue counters on this
  // device (vrt_pathtrace_noalloc), looked up once at creation
  uint32_t *static_queues = nullptr;
};

namespace vrt {

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int hip_fail(hipError_t e, const char *what);
int check_ctx(vrt_context ctx);

}  // namespace vrt

#define VRT_HIP(call)                                       \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return ::vrt::hip_fail(e_, #call); \
  } while (0)
