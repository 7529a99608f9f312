// C ABI of the neural radiance cache (include/vrt.h, vrt_nrc_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "abi_internal.h"
#include "nrc_internal.h"
#include "nrc_kernels.h"

namespace {

using namespace vrt;

void release_frame(vrt_nrc n) {
  for (void *p : {(void *)n->sample, (void *)n->query_of_pixel, (void *)n->q_raw, (void *)n->q_T,
                  (void *)n->q_out, (void *)n->t_raw, (void *)n->t_target, (void *)n->t_ratio,
                  (void *)n->t_query, (void *)n->ctr})
    if (p) (void)hipFree(p);
  n->sample = nullptr;
  n->query_of_pixel = nullptr;
  n->q_raw = n->q_T = n->q_out = n->t_raw = n->t_target = n->t_ratio = nullptr;
  n->t_query = nullptr;
  n->ctr = nullptr;
  n->pixels = 0;
  n->t_cap = 0;
}

void release(vrt_nrc n) {
  for (void *p : {(void *)n->d.master, (void *)n->d.adam_m, (void *)n->d.adam_v, (void *)n->d.w,
                  (void *)n->d.wt, (void *)n->d.bias, (void *)n->d.grad, (void *)n->d.partials,
                  (void *)n->d.stats})
    if (p) (void)hipFree(p);
  release_frame(n);
}

int check_nrc(vrt_nrc n) {
  if (!n) return fail(VRT_ERROR_INVALID_ARGUMENT, "null cache");
  return check_ctx(n->ctx);
}

}  // namespace

namespace vrt {

int nrc_frame_buffers(vrt_nrc n, size_t pixels, int t_cap) {
  if (pixels <= n->pixels && t_cap <= n->t_cap && n->ctr) return VRT_OK;
  VRT_HIP(hipStreamSynchronize(n->ctx->stream));
  const size_t P = std::max(pixels, n->pixels);
  const int T = std::max(t_cap, n->t_cap);
  release_frame(n);
  hipError_t e = hipSuccess;
  auto A = [&](void **p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 16));
  };
  A((void **)&n->sample, sizeof(float4) * P);
  A((void **)&n->query_of_pixel, sizeof(int) * P);
  A((void **)&n->q_raw, sizeof(float) * 6 * P);
  A((void **)&n->q_T, sizeof(float) * 3 * P);
  A((void **)&n->q_out, sizeof(float) * 3 * P);
  A((void **)&n->t_raw, sizeof(float) * 6 * size_t(T));
  A((void **)&n->t_target, sizeof(float) * 3 * size_t(T));
  A((void **)&n->t_ratio, sizeof(float) * 3 * size_t(T));
  A((void **)&n->t_query, sizeof(int) * size_t(T));
  A((void **)&n->ctr, sizeof(uint32_t) * 4);
  if (e == hipSuccess) e = hipMemsetAsync(n->ctr, 0, sizeof(uint32_t) * 4, n->ctx->stream);
  if (e != hipSuccess) {
    release_frame(n);
    return e == hipErrorOutOfMemory ? fail(VRT_ERROR_OUT_OF_MEMORY, "cache frame buffers")
                                    : hip_fail(e, "cache frame buffers");
  }
  n->pixels = P;
  n->t_cap = T;
  return VRT_OK;
}

int nrc_train_frame(vrt_nrc n, int steps) {
  for (int b = 0; b < steps; ++b) {
    VRT_HIP(nrc_launch_train_step(n->d, n->t_cap, n->t_raw, n->t_target, n->max_workgroups,
                                  n->cfg.learning_rate, n->cfg.beta1, n->cfg.beta2,
                                  n->cfg.epsilon, n->ctx->stream, &n->ctr[1], b, steps));
  }
  return VRT_OK;
}

}  // namespace vrt

extern "C" {

int vrt_nrc_param_count(void) { return kNrcParams; }

int vrt_nrc_create(vrt_context ctx, const vrt_nrc_config *cfg, vrt_nrc *out) {
  if (int r = check_ctx(ctx)) return r;
  if (!cfg || !out) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  if (cfg->precision != VRT_NRC_FP16 && cfg->precision != VRT_NRC_BF16)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "precision must be VRT_NRC_FP16 or VRT_NRC_BF16");
  if (cfg->max_batch < 1 || cfg->max_batch > (1 << 24))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "max_batch %d out of range", cfg->max_batch);
  auto *n = new vrt_nrc_s;
  n->ctx = ctx;
  n->cfg = *cfg;
  // one partial gradient per training workgroup (64 samples per chunk)
  n->max_workgroups = std::max(1, std::min((cfg->max_batch + 63) / 64, 2 * std::max(1, ctx->num_cus)));
  NrcDevice &d = n->d;
  d.bf16 = cfg->precision == VRT_NRC_BF16;
  hipError_t e = hipSuccess;
  auto A = [&](void **p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  A((void **)&d.master, sizeof(float) * kNrcParams);
  A((void **)&d.adam_m, sizeof(float) * kNrcParams);
  A((void **)&d.adam_v, sizeof(float) * kNrcParams);
  A((void **)&d.w, sizeof(uint16_t) * kNrcWeights);
  A((void **)&d.wt, sizeof(uint16_t) * kNrcWeights);
  A((void **)&d.bias, sizeof(float) * kNrcBiases);
  A((void **)&d.grad, sizeof(float) * kNrcParams);
  A((void **)&d.partials, sizeof(float) * kNrcPartStride * size_t(n->max_workgroups));
  A((void **)&d.stats, sizeof(float) * 4);
  if (e == hipSuccess) e = nrc_launch_init(d, cfg->seed, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    release(n);
    delete n;
    return hip_fail(e, "cache creation");
  }
  *out = n;
  return VRT_OK;
}

int vrt_nrc_destroy(vrt_nrc n) {
  if (!n) return VRT_OK;
  (void)hipSetDevice(n->ctx->device);
  (void)hipStreamSynchronize(n->ctx->stream);
  release(n);
  delete n;
  return VRT_OK;
}

int vrt_nrc_get_params(vrt_nrc n, float *params) {
  if (int r = check_nrc(n)) return r;
  if (!params) return fail(VRT_ERROR_INVALID_ARGUMENT, "null output");
  VRT_HIP(hipMemcpyAsync(params, n->d.master, sizeof(float) * kNrcParams, hipMemcpyDeviceToHost,
                         n->ctx->stream));
  VRT_HIP(hipStreamSynchronize(n->ctx->stream));
  return VRT_OK;
}

int vrt_nrc_set_params(vrt_nrc n, const float *params) {
  if (int r = check_nrc(n)) return r;
  if (!params) return fail(VRT_ERROR_INVALID_ARGUMENT, "null input");
  VRT_HIP(hipMemcpyAsync(n->d.master, params, sizeof(float) * kNrcParams, hipMemcpyHostToDevice,
                         n->ctx->stream));
  VRT_HIP(nrc_launch_refresh(n->d, n->ctx->stream));
  VRT_HIP(hipStreamSynchronize(n->ctx->stream));
  return VRT_OK;
}

int vrt_nrc_infer(vrt_nrc n, int count, const float *raw, float *rgb) {
  if (int r = check_nrc(n)) return r;
  if (count < 0 || (count > 0 && (!raw || !rgb)))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad inference arguments");
  VRT_HIP(nrc_launch_infer(n->d, count, raw, rgb, n->ctx->num_cus, n->ctx->stream));
  return VRT_OK;
}

int vrt_nrc_gradients(vrt_nrc n, int count, const float *raw, const float *target, float *grad,
                      float *loss) {
  if (int r = check_nrc(n)) return r;
  if (count < 0 || count > n->cfg.max_batch || (count > 0 && (!raw || !target)))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad gradient arguments (n=%d, max_batch=%d)", count,
                n->cfg.max_batch);
  VRT_HIP(nrc_launch_gradients(n->d, count, raw, target, n->max_workgroups, n->ctx->stream));
  if (grad)
    VRT_HIP(hipMemcpyAsync(grad, n->d.grad, sizeof(float) * kNrcParams, hipMemcpyDeviceToDevice,
                           n->ctx->stream));
  if (loss) {
    VRT_HIP(hipMemcpyAsync(loss, n->d.stats, sizeof(float), hipMemcpyDeviceToHost, n->ctx->stream));
    VRT_HIP(hipStreamSynchronize(n->ctx->stream));
  }
  return VRT_OK;
}

int vrt_nrc_train_step(vrt_nrc n, int count, const float *raw, const float *target, float *loss) {
  if (int r = check_nrc(n)) return r;
  if (count < 0 || count > n->cfg.max_batch || (count > 0 && (!raw || !target)))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad training arguments (n=%d, max_batch=%d)", count,
                n->cfg.max_batch);
  VRT_HIP(nrc_launch_train_step(n->d, count, raw, target, n->max_workgroups,
                                n->cfg.learning_rate, n->cfg.beta1, n->cfg.beta2,
                                n->cfg.epsilon, n->ctx->stream));
  if (loss) {
    VRT_HIP(hipMemcpyAsync(loss, n->d.stats, sizeof(float), hipMemcpyDeviceToHost, n->ctx->stream));
    VRT_HIP(hipStreamSynchronize(n->ctx->stream));
  }
  return VRT_OK;
}

int vrt_nrc_frame_counts(vrt_nrc n, uint32_t counts[4]) {
  if (int r = check_nrc(n)) return r;
  if (!counts) return fail(VRT_ERROR_INVALID_ARGUMENT, "null output");
  if (!n->ctr) {
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
    return VRT_OK;
  }
  VRT_HIP(hipMemcpyAsync(counts, n->ctr, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                         n->ctx->stream));
  VRT_HIP(hipStreamSynchronize(n->ctx->stream));
  return VRT_OK;
}

int vrt_mfma_probe(vrt_context ctx, const float *a, const float *b, float *c, int precision) {
  if (int r = check_ctx(ctx)) return r;
  if (!a || !b || !c) return fail(VRT_ERROR_INVALID_ARGUMENT, "null buffer");
  VRT_HIP(nrc_launch_mfma_probe(a, b, c, precision == VRT_NRC_BF16, ctx->stream));
  return VRT_OK;
}

}  // extern "C"
