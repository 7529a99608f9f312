// Neural radiance cache kernels (gfx950 MFMA): layout constants and launchers.
//
// Network: 121 inputs (5 scalars x 12 frequencies x {sin, cos} + density),
// 5 hidden layers of 64 ReLU units, 3 linear outputs.  Stored padded:
//   layer 0: 64 x 128   layers 1-4: 64 x 64   layer 5: 16 x 64
// weights row-major [out][in] (the MFMA A operand of Z^T = W X^T), plus a
// transposed copy [in][out] for the backward pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vrt {

constexpr int kNrcLayers = 6;
constexpr int kNrcInputs = 121;   // encoded features
constexpr int kNrcRaw = 6;        // raw inputs: position (3), direction (2), density (1)
constexpr int kNrcFreqs = 12;
constexpr int kNrcWidth = 64;

__host__ __device__ constexpr int nrc_in(int l) { return l == 0 ? 128 : 64; }
__host__ __device__ constexpr int nrc_out(int l) { return l == 5 ? 16 : 64; }
__host__ __device__ constexpr int nrc_woff(int l) { return l == 0 ? 0 : 8192 + (l - 1) * 4096; }
__host__ __device__ constexpr int nrc_boff(int l) { return l * 64; }
constexpr int kNrcWeights = 8192 + 4 * 4096 + 1024;  // 25600 padded weights
constexpr int kNrcBiases = 5 * 64 + 16;               // 336 padded biases
constexpr int kNrcParams = kNrcWeights + kNrcBiases;  // master parameter vector
// one workgroup's partial sums: the parameters' gradients, then its loss
constexpr int kNrcPartStride = (kNrcParams + 1 + 63) / 64 * 64;

// Parameter-vector index of weight (l, o, i) and bias (l, o).
__host__ __device__ inline int nrc_w(int l, int o, int i) { return nrc_woff(l) + o * nrc_in(l) + i; }
__host__ __device__ inline int nrc_b(int l, int o) { return kNrcWeights + nrc_boff(l) + o; }

struct NrcDevice {
  float *master;       // kNrcParams fp32 (weights then biases)
  float *adam_m;       // kNrcParams
  float *adam_v;       // kNrcParams
  uint16_t *w;         // kNrcWeights fp16/bf16 bits, [out][in]
  uint16_t *wt;        // kNrcWeights fp16/bf16 bits, transposed per layer [in][out]
  float *bias;         // kNrcBiases fp32 (added to the fp32 accumulators)
  float *grad;         // kNrcParams gradient sums of the last step
  float *partials;     // max_workgroups x kNrcPartStride
  float *stats;        // [0] loss sum of the last step, [1] samples, [2] Adam steps taken
  int bf16;            // 0: fp16, 1: bf16
};

hipError_t nrc_launch_init(const NrcDevice &d, uint64_t seed, hipStream_t s);
// n_dev (optional): a sample count written by an earlier kernel; the
// kernels then use min(n, *n_dev) samples, so no host round trip is needed
hipError_t nrc_launch_infer(const NrcDevice &d, int n, const float *raw, float *out,
                            int num_cus, hipStream_t s, const uint32_t *n_dev = nullptr);
// gradient sums of the relative-L2 loss over mini-batch `batch` of
// `nbatches` into d.grad (and the loss sum into d.stats[0]); does not update
// the parameters
hipError_t nrc_launch_gradients(const NrcDevice &d, int n, const float *raw,
                                const float *target, int max_workgroups, hipStream_t s,
                                const uint32_t *n_dev = nullptr, int batch = 0,
                                int nbatches = 1);
// the same gradients, then one Adam step on the mini-batch's mean gradient;
// an empty mini-batch changes nothing, not even the step count of the bias
// correction
hipError_t nrc_launch_train_step(const NrcDevice &d, int n, const float *raw,
                                 const float *target, int max_workgroups, float lr, float beta1,
                                 float beta2, float eps, hipStream_t s,
                                 const uint32_t *n_dev = nullptr, int batch = 0,
                                 int nbatches = 1);
hipError_t nrc_launch_refresh(const NrcDevice &d, hipStream_t s);

// MFMA lane-layout probe: C = A (16x16) * B (16x16) for one wave.
hipError_t nrc_launch_mfma_probe(const float *a, const float *b, float *c, int bf16,
                                 hipStream_t s);

}  // namespace vrt
