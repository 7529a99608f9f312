// vrt_hdri: the HDRI light's map and sampling tables on the device
// (HDRI.cpp:31-162).  The CDFs are built on the host as makeCDF builds them:
// per row an inclusive running sum of the max-component luminance, divided by
// the row's total (rows summing to 0 are left as they are); the totals' own
// running sum, divided by its last entry, selects the row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "abi_internal.h"
#include "hdri_internal.h"
#include "vrt.h"

namespace vrt {
namespace {

void scan_normalize(float *d, size_t n) {
  for (size_t i = 1; i < n; ++i) d[i] = d[i - 1] + d[i];
  const float last = d[n - 1];
  if (last != 0.f)
    for (size_t i = 0; i < n; ++i) d[i] /= last;
}

f3 norm3(f3 a) { return a * (1.f / length(a)); }

}  // namespace

bool hdri_frame(const float up[3], const float direction[3], float W[9], float L[9]) {
  const f3 c1 = -norm3(mk(up[0], up[1], up[2]));
  const f3 c0 = norm3(cross(c1, mk(direction[0], direction[1], direction[2])));
  const f3 c2 = norm3(cross(c0, c1));
  const f3 cols[3] = {c0, c1, c2};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) W[3 * r + c] = comp(cols[c], r);
  // inverse by the adjugate
  const float det = W[0] * (W[4] * W[8] - W[5] * W[7]) - W[1] * (W[3] * W[8] - W[5] * W[6]) +
                    W[2] * (W[3] * W[7] - W[4] * W[6]);
  if (!(det != 0.f) || !std::isfinite(det)) return false;
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
  return true;
}

}  // namespace vrt

using namespace vrt;

extern "C" {

int vrt_hdri_create(vrt_context ctx, const float *radiance, int w, int h, vrt_hdri *out) {
  if (int r = check_ctx(ctx)) return r;
  if (!out || !radiance) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (w < 1 || h < 1 || int64_t(w) * h > (int64_t(1) << 28))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "hdri size %dx%d out of range", w, h);
  const size_t n = size_t(w) * h;
  std::vector<float4> rgba(n);
  std::vector<float> cols(n), rows(h);
  for (size_t i = 0; i < n; ++i) {
    const float r = radiance[3 * i], g = radiance[3 * i + 1], b = radiance[3 * i + 2];
    rgba[i] = make_float4(r, g, b, 1.f);
    cols[i] = std::max(std::max(r, g), b);  // luminance as makeCDF takes it
  }
  for (int y = 0; y < h; ++y) {
    float *row = cols.data() + size_t(y) * w;
    for (int x = 1; x < w; ++x) row[x] = row[x - 1] + row[x];
    rows[y] = row[w - 1];
    const float last = row[w - 1];
    if (last != 0.f)
      for (int x = 0; x < w; ++x) row[x] /= last;
  }
  scan_normalize(rows.data(), rows.size());
  auto *d = new vrt_hdri_s;
  d->device = ctx->device;
  d->w = w;
  d->h = h;
  hipError_t e = hipMalloc(&d->tex, n * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&d->rows, size_t(h) * 4);
  if (e == hipSuccess) e = hipMalloc(&d->cols, n * 4);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d->tex, rgba.data(), n * sizeof(float4), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d->rows, rows.data(), size_t(h) * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d->cols, cols.data(), n * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // host copies go out of scope
  if (e != hipSuccess) {
    vrt_hdri_destroy(d);
    return hip_fail(e, "hdri upload");
  }
  *out = d;
  return VRT_OK;
}

int vrt_hdri_destroy(vrt_hdri d) {
  if (!d) return VRT_OK;
  (void)hipSetDevice(d->device);
  (void)hipFree(d->tex);
  (void)hipFree(d->rows);
  (void)hipFree(d->cols);
  delete d;
  return VRT_OK;
}

}  // extern "C"
