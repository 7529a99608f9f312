// Neural radiance cache on gfx950 matrix cores.
//
// Every layer is evaluated transposed, Z^T (out x M) = W (out x in) X^T
// (in x M), with gfx950's v_mfma_f32_16x16x32 (fp16 or bf16 operands, fp32
// accumulation: twice the work of the 16x16x16 form per instruction).  Lane
// l of a wave holds, for one 16 x 32 x 16 step,
//   A operand: W[l%16][8*(l/16) + j]        (j = 0..7)
//   B operand: X^T[8*(l/16) + j][l%16]
//   D result : Z^T[4*(l/16) + j][l%16]      (j = 0..3)
// A lane's results of output blocks 2b and 2b+1 (rows 4g..4g+3 of each,
// g = l/16), after bias + ReLU and rounding to 16 bits, *are* its B operand
// of input block b of the next layer, provided that layer reads its inputs
// in the order kperm below: activations stay in registers from the encoding
// to the output.  The 16-bit weight copies are stored with their input
// columns in that order (layers 1-5), so each lane reads its 8 weights with
// one 16-byte load.
//
// Training keeps the same per-wave forward (16 samples per wave, 64 per
// workgroup), stores each layer's input X^T and the gradient dZ^T in LDS
// ([feature][sample], 16-bit), back-propagates dX^T = W^T dZ^T per wave in
// registers, and forms dW = dZ^T X over the workgroup's 64 samples with one
// MFMA chain per 16x16 weight tile (tiles dealt round-robin to the 4 waves,
// accumulated in registers across the workgroup's sample chunks).  A
// training step is latency-bound (one 64-sample chunk per workgroup at 8,192
// samples), so the weights of layers 1-5 and their transposes are copied to
// LDS with one round of loads at the start rather than fetched layer by
// layer.  Partial gradients and losses per workgroup are summed in a fixed
// order (deterministic, bias sums included), and the same kernel applies
// Adam to the fp32 master weights and re-emits the 16-bit copies.
#include "nrc_kernels.h"

#include <algorithm>

namespace vrt {

typedef _Float16 h4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// 16-bit conversions, both IEEE round-to-nearest-even in hardware (gfx950
// converts fp32 to bf16 with v_cvt_pk_bf16_f32)
// ---------------------------------------------------------------------------
__device__ inline uint16_t to16(float f, int bf16) {
  if (bf16) return __builtin_bit_cast(uint16_t, (__bf16)f);
  _Float16 h = (_Float16)f;
  return __builtin_bit_cast(uint16_t, h);
}
__device__ inline float from16(uint16_t b, int bf16) {
  if (bf16) return __uint_as_float(uint32_t(b) << 16);
  return (float)__builtin_bit_cast(_Float16, b);
}

struct Frag {  // 4 x 16-bit operand of one lane
  uint2 v;
};
__device__ inline Frag pack4(float a, float b, float c, float d, int bf16) {
  Frag f;
  f.v.x = uint32_t(to16(a, bf16)) | (uint32_t(to16(b, bf16)) << 16);
  f.v.y = uint32_t(to16(c, bf16)) | (uint32_t(to16(d, bf16)) << 16);
  return f;
}
__device__ inline float unpack(const Frag &f, int j, int bf16) {
  const uint32_t w = j < 2 ? f.v.x : f.v.y;
  return from16(uint16_t((j & 1) ? (w >> 16) : (w & 0xFFFFu)), bf16);
}

struct Frag8 {  // 8 x 16-bit operand of one lane (16x16x32 MFMA)
  uint4 v;
};
__device__ inline Frag8 cat(const Frag &lo, const Frag &hi) {
  Frag8 f;
  f.v = make_uint4(lo.v.x, lo.v.y, hi.v.x, hi.v.y);
  return f;
}
__device__ inline Frag8 load8(const uint16_t *p) {
  Frag8 f;
  f.v = *reinterpret_cast<const uint4 *>(p);
  return f;
}

// 16x16x16: the output layer's backward pass (16 outputs)
template <int BF16>
__device__ inline f4v mfma(const Frag &a, const Frag &b, f4v c) {
  if constexpr (BF16) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4v, a.v),
                                                     __builtin_bit_cast(s4v, b.v), c, 0, 0, 0);
  } else {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(h4v, a.v),
                                                 __builtin_bit_cast(h4v, b.v), c, 0, 0, 0);
  }
}
// 16x16x32: everything else
template <int BF16>
__device__ inline f4v mfma(const Frag8 &a, const Frag8 &b, f4v c) {
  if constexpr (BF16) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8v, a.v),
                                                   __builtin_bit_cast(b8v, b.v), c, 0, 0, 0);
  } else {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8v, a.v),
                                                  __builtin_bit_cast(h8v, b.v), c, 0, 0, 0);
  }
}

// Order in which a layer fed by a previous layer's results reads its 64
// inputs: in each 32-input block b, position 8g + j holds input
// 32b + 4g + j (j < 4) or 32b + 16 + 4g + (j - 4) -- the rows lane group g
// holds of output blocks 2b and 2b+1.  kperm maps an input index to its
// position.
__host__ __device__ constexpr int kperm(int c) {
  return (c & ~31) + (((c & 15) >> 2) << 3) + (((c >> 4) & 1) << 2) + (c & 3);
}
__host__ __device__ constexpr int weight_col(int l, int c) { return l == 0 ? c : kperm(c); }

// ---------------------------------------------------------------------------
// Input encoding: feature f < 120 is sin/cos(2^k pi x_s) with s = f / 24,
// k = (f % 24) / 2, sin for even f; f = 120 is the density; 121..127 pad.
// v_sin/v_cos take revolutions: 2^k pi x = 2 pi * frac(2^(k-1) x), and the
// scaling by a power of two and the fraction are exact.
// ---------------------------------------------------------------------------
// Raw input of a lane-dependent feature f < 120 (input f / 24), as
// comparisons on f: an indexed read of the 6-element array would go through
// scratch or movrel.
__device__ inline float raw_at(const float r[6], int f) {
  return f < 48 ? (f < 24 ? r[0] : r[1]) : (f < 72 ? r[2] : (f < 96 ? r[3] : r[4]));
}

// Features f0 .. f0 + 7 (f0 a multiple of 8, so one input's frequencies k0 ..
// k0 + 3, or the density and padding): sin/cos of the first frequency from
// the hardware, the next three by the double-angle identities sin 2a =
// 2 sin a cos a, cos 2a = 1 - 2 sin^2 a.  Three doublings scale the
// first pair's rounding error by at most 8 (a few 1e-7), far below the 16-bit
// rounding of the features.
__device__ inline void encode8(const float *raw, int f0, float e[8]) {
  if (f0 < 120) {
    const int k0 = (f0 % 24) >> 1;
    float r = raw_at(raw, f0) * __uint_as_float(uint32_t(126 + k0) << 23);  // * 2^(k0-1)
    r = r - floorf(r);
    float sn = __builtin_amdgcn_sinf(r), cs = __builtin_amdgcn_cosf(r);
    e[0] = sn;
    e[1] = cs;
#pragma unroll
    for (int i = 1; i < 4; ++i) {
      const float s2 = 2.f * sn * cs;
      cs = 1.f - 2.f * sn * sn;
      sn = s2;
      e[2 * i] = sn;
      e[2 * i + 1] = cs;
    }
  } else {
    e[0] = f0 == 120 ? raw[5] : 0.f;
#pragma unroll
    for (int j = 1; j < 8; ++j) e[j] = 0.f;
  }
}

__device__ inline void load_raw(const float *raw, int n, int q, float r[6]) {
  if (q < n) {
#pragma unroll
    for (int i = 0; i < 6; ++i) r[i] = raw[6 * size_t(q) + i];
  } else {
#pragma unroll
    for (int i = 0; i < 6; ++i) r[i] = 0.f;
  }
}

// ---------------------------------------------------------------------------
// Forward pass of QB 16-query blocks per wave.  `hin` holds the layer-0 B
// operands (4 k-blocks of 32 features); the 3 outputs of query block qb are
// returned in out[qb] (rows 0..3 of lanes 0..15).  Weights come through
// `w0frag(ob, kb)` -> this lane's A operand of layer 0's output block ob and
// input block kb; `wrow(l, row, col)` (layers 1-5) -> pointer to 8
// consecutive 16-bit weights of the stored (kperm-ordered) row.
// ---------------------------------------------------------------------------
template <int BF16, int QB, class W0Frag, class WPtr, class Store>
__device__ inline void forward(const W0Frag &w0frag, const WPtr &wrow, const float *bias,
                               Frag8 (&hin)[4][QB], f4v (&out)[QB], const Store &store_act) {
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15, g = lane >> 4;
  Frag h[4][QB];
  // layer 0: 128 -> 64
  {
    f4v acc[4][QB];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
      for (int qb = 0; qb < QB; ++qb)
        for (int j = 0; j < 4; ++j) acc[ob][qb][j] = bias[nrc_boff(0) + ob * 16 + 4 * g + j];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) {
        const Frag8 a = w0frag(ob, kb);
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) acc[ob][qb] = mfma<BF16>(a, hin[kb][qb], acc[ob][qb]);
      }
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        h[ob][qb] = pack4(fmaxf(acc[ob][qb][0], 0.f), fmaxf(acc[ob][qb][1], 0.f),
                          fmaxf(acc[ob][qb][2], 0.f), fmaxf(acc[ob][qb][3], 0.f), BF16);
        store_act(1, ob, qb, h[ob][qb]);
      }
  }
  // hidden layers 1..4: 64 -> 64; input block kb is output blocks 2kb, 2kb+1
#pragma unroll 1
  for (int l = 1; l <= 4; ++l) {
    f4v acc[4][QB];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
      for (int qb = 0; qb < QB; ++qb)
        for (int j = 0; j < 4; ++j) acc[ob][qb][j] = bias[nrc_boff(l) + ob * 16 + 4 * g + j];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) {
        const Frag8 a = load8(wrow(l, ob * 16 + r16, kb * 32 + 8 * g));
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
          acc[ob][qb] = mfma<BF16>(a, cat(h[2 * kb][qb], h[2 * kb + 1][qb]), acc[ob][qb]);
      }
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        h[ob][qb] = pack4(fmaxf(acc[ob][qb][0], 0.f), fmaxf(acc[ob][qb][1], 0.f),
                          fmaxf(acc[ob][qb][2], 0.f), fmaxf(acc[ob][qb][3], 0.f), BF16);
        store_act(l + 1, ob, qb, h[ob][qb]);
      }
  }
  // output layer 5: 64 -> 16 (3 used), linear
#pragma unroll
  for (int qb = 0; qb < QB; ++qb)
    for (int j = 0; j < 4; ++j) out[qb][j] = bias[nrc_boff(5) + 4 * g + j];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    const Frag8 a = load8(wrow(5, r16, kb * 32 + 8 * g));
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
      out[qb] = mfma<BF16>(a, cat(h[2 * kb][qb], h[2 * kb + 1][qb]), out[qb]);
  }
}

// ---------------------------------------------------------------------------
// Inference: persistent workgroups, weights staged in LDS with rows padded by
// 8 halves (row stride 16 B off a multiple of 128 B: the 16 rows one
// ds_read_b128 phase touches spread over all banks); 16 * kInferQB queries
// per wave.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int lds_stride(int l) { return nrc_in(l) + 8; }
__host__ __device__ constexpr int lds_woff(int l) {
  return l == 0 ? 0 : 64 * lds_stride(0) + (l - 1) * 64 * lds_stride(1);
}
constexpr int kLdsWeights = 64 * (128 + 8) + 4 * 64 * (64 + 8) + 16 * (64 + 8);
// 16-query blocks per wave in inference: 3 keeps the kernel at two waves per
// SIMD (163 VGPRs + 84 AGPRs; 4 blocks need 167 + 112 and fall to one wave),
// which the 64 KB of staged weights allow per CU anyway
constexpr int kInferQB = 3;

// Sample range [lo, lo + cnt) of mini-batch `batch` of `nbatches` over the
// first min(n, *n_dev) samples (n_dev: a count produced on the GPU, or null).
__device__ inline void batch_range(int n, const uint32_t *n_dev, int batch, int nbatches, int &lo,
                                   int &cnt) {
  const int total = n_dev ? min(n, int(*n_dev)) : n;
  lo = int((long long)total * batch / nbatches);
  cnt = int((long long)total * (batch + 1) / nbatches) - lo;
}

template <int BF16>
__global__ void __launch_bounds__(256) k_nrc_infer(NrcDevice d, int n, const uint32_t *n_dev,
                                                   const float *__restrict__ raw,
                                                   float *__restrict__ out) {
  {
    int lo;
    batch_range(n, n_dev, 0, 1, lo, n);
  }
  __shared__ __align__(16) uint16_t sw[kLdsWeights];
  __shared__ float sb[kNrcBiases];
  for (int l = 0; l < kNrcLayers; ++l) {
    const int rows = nrc_out(l), cols = nrc_in(l);
    for (int i = threadIdx.x; i < rows * cols; i += blockDim.x)
      sw[lds_woff(l) + (i / cols) * lds_stride(l) + (i % cols)] = d.w[nrc_woff(l) + i];
  }
  for (int i = threadIdx.x; i < kNrcBiases; i += blockDim.x) sb[i] = d.bias[i];
  __syncthreads();

  auto wrow = [&](int l, int row, int col) -> const uint16_t * {
    return &sw[lds_woff(l) + row * lds_stride(l) + col];
  };
  auto no_store = [](int, int, int, const Frag &) {};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, g = lane >> 4;
  auto w0frag = [&](int ob, int kb) { return load8(wrow(0, ob * 16 + r16, kb * 32 + 8 * g)); };
  constexpr int QB = kInferQB;
  constexpr int TQ = 16 * QB;  // queries per wave-tile
  const int tiles = (n + TQ - 1) / TQ;
  for (int t = blockIdx.x * 4 + wave; t < tiles; t += gridDim.x * 4) {
    Frag8 hin[4][QB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      float r[6];
      load_raw(raw, n, t * TQ + qb * 16 + r16, r);
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        const int f = kb * 32 + 8 * g;
        float e[8];
        encode8(r, f, e);
        hin[kb][qb] = cat(pack4(e[0], e[1], e[2], e[3], BF16), pack4(e[4], e[5], e[6], e[7], BF16));
      }
    }
    f4v o[QB];
    forward<BF16, QB>(w0frag, wrow, sb, hin, o, no_store);
    if (g == 0) {
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        const int q = t * TQ + qb * 16 + r16;
        if (q < n) {
          out[3 * size_t(q)] = o[qb][0];
          out[3 * size_t(q) + 1] = o[qb][1];
          out[3 * size_t(q) + 2] = o[qb][2];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Training gradients (relative L2 loss, Mueller et al. 2021):
//   L = sum_q sum_c (p - t)^2 / (p^2 + 0.01),  dL/dp = 2 (p - t) / (p^2 + 0.01)
// with p^2 in the denominator treated as a constant.
// ---------------------------------------------------------------------------
constexpr int kTS = 64 + 8;  // LDS row stride of [feature][sample] tiles (halves, 16-B rows)
// The gradient kernel's LDS copies of W and W^T of layers 1-5: rows of 64
// halves padded to 72, except W^T of layer 5 (rows of 16 outputs, padded to
// 24).  Layer l starts at (l - 1) * 64 * 72 in either copy.
__host__ __device__ constexpr int gw_off(int l) { return (l - 1) * 64 * kTS; }
__host__ __device__ constexpr int gwt_stride(int l) { return l == 5 ? 24 : kTS; }
constexpr int kGradW = 4 * 64 * kTS + 16 * kTS;
constexpr int kGradWt = 4 * 64 * kTS + 64 * 24;

// Sum over the 16 lanes of a DPP row, the same addition tree as butterfly
// exchanges with lane distance 1, 2, 4, 8 (addition is commutative, so every
// lane gets the bit-identical sum): quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror, row_mirror -- no LDS round trip.
__device__ inline float row_sum16(float s) {
  s += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0xB1, 0xF, 0xF, false));
  s += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0x4E, 0xF, 0xF, false));
  s += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0x141, 0xF, 0xF, false));
  s += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0x140, 0xF, 0xF, false));
  return s;
}

//
// count_step: the launch belongs to an Adam step, whose bias-correction step
// count (d.stats[2]) advances here when the mini-batch holds samples (a
// frame's sample count is known only on the device); nothing in this kernel
// reads it, the update kernel that follows does.
template <int BF16>
__global__ void __launch_bounds__(256) k_nrc_grad(NrcDevice d, int n, const uint32_t *n_dev,
                                                  int batch, int nbatches, int count_step,
                                                  const float *__restrict__ raw,
                                                  const float *__restrict__ target) {
  {
    int lo;
    batch_range(n, n_dev, batch, nbatches, lo, n);
    raw += 6 * size_t(lo);
    target += 3 * size_t(lo);
  }
  if (count_step && n > 0 && blockIdx.x == 0 && threadIdx.x == 0) d.stats[2] += 1.f;
  __shared__ __align__(16) uint16_t sx0[128 * kTS];   // layer-0 input X^T [feature][sample]
  __shared__ __align__(16) uint16_t sx[5][64 * kTS];  // inputs of layers 1..5
  __shared__ __align__(16) uint16_t sdz[64 * kTS];    // dZ^T of the current layer [out][sample]
  __shared__ float sdb[4][kNrcBiases];     // bias-gradient sums of each wave
  __shared__ float sloss[4];               // loss sums of each wave
  // W and W^T of layers 1-5 and the biases: every layer's weights arrive
  // with one round of loads at the start instead of one dependent global
  // round trip per layer and direction (layer 0's W, 16 KB, is read from
  // global memory by the forward pass)
  __shared__ __align__(16) uint16_t sw[kGradW];
  __shared__ __align__(16) uint16_t swt[kGradWt];
  __shared__ __align__(16) float sbias[kNrcBiases];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, g = lane >> 4;
  const int col = wave * 16 + r16;  // this lane's sample column in a chunk
  // the first chunk's inputs and layer 0's weights (this lane's A operands,
  // held in registers) are requested together with the LDS copies below
  float r[6];
  load_raw(raw, n, blockIdx.x * 64 + col, r);
  Frag8 w0[4][4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
      w0[kb][ob] = load8(&d.w[nrc_woff(0) + (ob * 16 + r16) * nrc_in(0) + kb * 32 + 8 * g]);
  // W and W^T of layers 1-4: 64 rows x 8 chunks of 8 halves each, 8
  // unconditional chunks of either per thread; layer 5: W 16 rows x 8
  // chunks (threads 0-127), W^T 64 rows x 2 chunks (threads 128-255); the
  // biases as 84 float4.  Every load is issued before the first LDS store.
  const int j5 = threadIdx.x & 127;
  const bool tr5 = threadIdx.x >= 128;
  const uint4 v5 = *reinterpret_cast<const uint4 *>(
      tr5 ? &d.wt[nrc_woff(5) + (j5 >> 1) * 16 + 8 * (j5 & 1)]
          : &d.w[nrc_woff(5) + (j5 >> 3) * 64 + 8 * (j5 & 7)]);
  const float4 vb =
      reinterpret_cast<const float4 *>(d.bias)[min(int(threadIdx.x), kNrcBiases / 4 - 1)];
  uint4 vw[8], vt[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = threadIdx.x + 256 * k;
    const int l = 1 + (i >> 9), r = (i >> 3) & 63, c = i & 7;
    vw[k] = *reinterpret_cast<const uint4 *>(&d.w[nrc_woff(l) + r * 64 + 8 * c]);
    vt[k] = *reinterpret_cast<const uint4 *>(&d.wt[nrc_woff(l) + r * 64 + 8 * c]);
  }
  // the encoding of chunk c's inputs r: layer 0's B operands in registers,
  // X^T of layer 0 in LDS
  auto encode_chunk = [&](int c, Frag8 (&hin)[4][1]) {
    const bool valid = c * 64 + col < n;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      const int f = kb * 32 + 8 * g;
      float e[8];
      encode8(r, f, e);
      for (int j = 0; j < 8; ++j) e[j] = valid ? e[j] : 0.f;
      hin[kb][0] = cat(pack4(e[0], e[1], e[2], e[3], BF16), pack4(e[4], e[5], e[6], e[7], BF16));
      for (int j = 0; j < 8; ++j) sx0[(f + j) * kTS + col] = to16(e[j], BF16);
    }
  };
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = threadIdx.x + 256 * k;
    const int l = 1 + (i >> 9), r = (i >> 3) & 63, c = i & 7;
    *reinterpret_cast<uint4 *>(&sw[gw_off(l) + r * kTS + 8 * c]) = vw[k];
    *reinterpret_cast<uint4 *>(&swt[gw_off(l) + r * kTS + 8 * c]) = vt[k];
  }
  *reinterpret_cast<uint4 *>(tr5 ? &swt[gw_off(5) + (j5 >> 1) * 24 + 8 * (j5 & 1)]
                                 : &sw[gw_off(5) + (j5 >> 3) * kTS + 8 * (j5 & 7)]) = v5;
  if (threadIdx.x < kNrcBiases / 4) reinterpret_cast<float4 *>(sbias)[threadIdx.x] = vb;
  for (int i = threadIdx.x; i < 4 * kNrcBiases; i += blockDim.x) (&sdb[0][0])[i] = 0.f;
  if (threadIdx.x < 4) sloss[threadIdx.x] = 0.f;
  __syncthreads();

  // dW accumulators: tiles of each layer dealt round-robin to the waves
  // layer 0: 4x8 tiles (8 per wave), layers 1-4: 4x4 (4 per wave), layer 5:
  // 1x4 (1 per wave)
  f4v acc0[8], acc14[4][4], acc5;
  for (int i = 0; i < 8; ++i) acc0[i] = f4v{0.f, 0.f, 0.f, 0.f};
  for (int l = 0; l < 4; ++l)
    for (int i = 0; i < 4; ++i) acc14[l][i] = f4v{0.f, 0.f, 0.f, 0.f};
  acc5 = f4v{0.f, 0.f, 0.f, 0.f};

  auto w0frag = [&](int ob, int kb) { return w0[kb][ob]; };
  auto wrow = [&](int l, int row, int col) -> const uint16_t * {
    return &sw[gw_off(l) + row * kTS + col];
  };
  auto store_act = [&](int l, int ob, int, const Frag &h) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t w = j < 2 ? h.v.x : h.v.y;
      sx[l - 1][(ob * 16 + 4 * g + j) * kTS + col] = uint16_t((j & 1) ? (w >> 16) : (w & 0xFFFFu));
    }
  };

  const int chunks = (n + 63) / 64;
  for (int c = blockIdx.x; c < chunks; c += gridDim.x) {
    const int q = c * 64 + col;
    const bool valid = q < n;
    if (c != int(blockIdx.x)) load_raw(raw, n, q, r);
    Frag8 hin[4][1];
    encode_chunk(c, hin);
    f4v out[1];
    forward<BF16, 1>(w0frag, wrow, sbias, hin, out, store_act);

    // output gradient: rows 0..2 of lanes 0..15 hold the prediction
    float gz[4] = {0.f, 0.f, 0.f, 0.f};
    float ls = 0.f;
    if (g == 0 && valid) {
      for (int ch = 0; ch < 3; ++ch) {
        const float p = out[0][ch], t = target[3 * size_t(q) + ch];
        const float den = p * p + 0.01f;
        const float diff = p - t;
        ls += diff * diff / den;
        gz[ch] = 2.f * diff / den;
      }
    }
    ls = row_sum16(ls);
    if (lane == 0) sloss[wave] += ls;
    // ---- backward ------------------------------------------------------------
    // dZ^T of the current layer for this wave's 16 samples, D layout,
    // [ob][j] -> row ob*16 + 4g + j
    float dz[4][4];
    for (int ob = 0; ob < 4; ++ob)
      for (int j = 0; j < 4; ++j) dz[ob][j] = 0.f;
    for (int j = 0; j < 4; ++j) dz[0][j] = gz[j];
    __syncthreads();  // forward activations of all waves are in LDS

#pragma unroll
    for (int l = 5; l >= 0; --l) {
      const int nob = nrc_out(l) / 16, nib = nrc_in(l) / 16;
      // stage dZ^T, accumulate the bias gradient
      Frag dzf[4];
      for (int ob = 0; ob < nob; ++ob) {
        dzf[ob] = pack4(dz[ob][0], dz[ob][1], dz[ob][2], dz[ob][3], BF16);
        for (int j = 0; j < 4; ++j) {
          sdz[(ob * 16 + 4 * g + j) * kTS + col] = to16(dz[ob][j], BF16);
          // one lane per entry of this wave's sums: the additions happen in
          // program order, and an atomic add returns nothing to wait for
          const float s = row_sum16(dz[ob][j]);
          if (r16 == 0) atomicAdd(&sdb[wave][nrc_boff(l) + ob * 16 + 4 * g + j], s);
        }
      }
      __syncthreads();
      // dW tiles (ob, ib) over the chunk's 64 samples (k = sample)
      const uint16_t *xin = l == 0 ? sx0 : sx[l - 1];
      // nob * nib is a multiple of 4: every wave owns the same number of
      // tiles, and with l and k known at compile time the accumulators stay
      // in registers
#pragma unroll
      for (int k = 0; k < nob * nib / 4; ++k) {
        const int t = wave + 4 * k;
        const int ob = t / nib, ib = t % nib;
        f4v acc = l == 0 ? acc0[k & 7] : (l == 5 ? acc5 : acc14[(l - 1) & 3][k & 3]);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
          const Frag8 a = load8(&sdz[(ob * 16 + r16) * kTS + kb * 32 + 8 * g]);
          const Frag8 b = load8(&xin[(ib * 16 + r16) * kTS + kb * 32 + 8 * g]);
          acc = mfma<BF16>(a, b, acc);
        }
        if (l == 0)
          acc0[k & 7] = acc;
        else if (l == 5)
          acc5 = acc;
        else
          acc14[(l - 1) & 3][k & 3] = acc;
      }
      if (l > 0) {
        // dX^T = W^T dZ^T (k = output of layer l), then the ReLU mask of
        // layer l-1's output (= X of l).  64 outputs: two 32-blocks read in
        // kperm order; the output layer's 16: one 16x16x16 step
        for (int ib = 0; ib < nib; ++ib) {
          f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
          const uint16_t *wt = &swt[gw_off(l) + (ib * 16 + r16) * gwt_stride(l)];
          if (nob == 1) {
            Frag a;
            a.v = *reinterpret_cast<const uint2 *>(wt + 4 * g);
            acc = mfma<BF16>(a, dzf[0], acc);
          } else {
            for (int kb = 0; kb < 2; ++kb)
              acc = mfma<BF16>(load8(wt + kb * 32 + 8 * g), cat(dzf[2 * kb], dzf[2 * kb + 1]), acc);
          }
          for (int j = 0; j < 4; ++j) {
            const float x = from16(xin[(ib * 16 + 4 * g + j) * kTS + col], BF16);
            dz[ib][j] = x > 0.f ? acc[j] : 0.f;
          }
        }
      }
      __syncthreads();  // sdz is rewritten by the next layer
    }
  }

  // ---- write this workgroup's partial gradient and loss -----------------------
  float *part = d.partials + size_t(blockIdx.x) * kNrcPartStride;
  auto put = [&](int l, int t, const f4v &acc) {
    const int nib = nrc_in(l) / 16;
    const int ob = t / nib, ib = t % nib;
    for (int j = 0; j < 4; ++j) part[nrc_w(l, ob * 16 + 4 * g + j, ib * 16 + r16)] = acc[j];
  };
#pragma unroll
  for (int k = 0; k < 8; ++k) put(0, wave + 4 * k, acc0[k]);
#pragma unroll
  for (int l = 1; l <= 4; ++l)
#pragma unroll
    for (int k = 0; k < 4; ++k) put(l, wave + 4 * k, acc14[l - 1][k]);
  put(5, wave, acc5);
  __syncthreads();
  for (int i = threadIdx.x; i < kNrcBiases; i += blockDim.x)
    part[kNrcWeights + i] = (sdb[0][i] + sdb[1][i]) + (sdb[2][i] + sdb[3][i]);
  if (threadIdx.x == 0) part[kNrcParams] = (sloss[0] + sloss[1]) + (sloss[2] + sloss[3]);
}

__device__ inline void emit16(const NrcDevice &d, int p, float v) {
  if (p < kNrcWeights) {
    int l = 5;
    while (l > 0 && p < nrc_woff(l)) --l;
    const int local = p - nrc_woff(l);
    const int o = local / nrc_in(l), i = local % nrc_in(l);
    const uint16_t b = to16(v, d.bf16);
    // W [out][in], inputs of layers 1-5 in kperm order; W^T [in][out],
    // outputs in kperm order where they are 64 (layers 1-4)
    d.w[nrc_woff(l) + o * nrc_in(l) + weight_col(l, i)] = b;
    d.wt[nrc_woff(l) + i * nrc_out(l) + (nrc_out(l) == 64 ? kperm(o) : o)] = b;
  } else {
    d.bias[p - kNrcWeights] = v;
  }
}

// Fixed-order sum of the workgroup partials: a workgroup owns 64 entries
// (parameters, and at kNrcParams the loss); its 4 waves sum interleaved
// quarters of the partials (64 consecutive floats per load), then combine in
// LDS in a fixed order.  With ADAM the sum is the gradient of an Adam step
// (Kingma & Ba 2015, bias-corrected) on the fp32 master parameters, which
// then re-emits the 16-bit copies; the step count was advanced by the
// gradient launch, and an empty mini-batch changes nothing.
template <int ADAM>
__global__ void __launch_bounds__(256) k_nrc_reduce(NrcDevice d, int parts, int n,
                                                    const uint32_t *n_dev, int batch,
                                                    int nbatches, float lr, float b1, float b2,
                                                    float eps) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + lane;
  float s = 0.f;
  if (p <= kNrcParams) {
    // batches of 16 loads in flight, summed in partial order
    const float *src = d.partials + p;
    int i = wave;
    for (; i + 4 * 15 < parts; i += 4 * 16) {
      float v[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = src[size_t(i + 4 * k) * kNrcPartStride];
#pragma unroll
      for (int k = 0; k < 16; ++k) s += v[k];
    }
    for (; i < parts; i += 4) s += src[size_t(i) * kNrcPartStride];
  }
  __shared__ float sh[4][64];
  sh[wave][lane] = s;
  __syncthreads();
  if (wave != 0 || p > kNrcParams) return;
  const float sum = (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]);
  if (p == kNrcParams) {
    d.stats[0] = sum;
    return;
  }
  d.grad[p] = sum;
  if (!ADAM) return;
  int lo, cnt;
  batch_range(n, n_dev, batch, nbatches, lo, cnt);
  if (cnt <= 0) return;
  const float step = d.stats[2];
  const float c1 = 1.f / (1.f - powf(b1, step));
  const float c2 = 1.f / (1.f - powf(b2, step));
  const float gr = sum * (1.f / float(cnt));
  const float m = b1 * d.adam_m[p] + (1.f - b1) * gr;
  const float v = b2 * d.adam_v[p] + (1.f - b2) * gr * gr;
  d.adam_m[p] = m;
  d.adam_v[p] = v;
  const float w = d.master[p] - lr * (m * c1) / (sqrtf(v * c2) + eps);
  d.master[p] = w;
  emit16(d, p, w);
}

__global__ void k_nrc_refresh(NrcDevice d) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNrcParams) return;
  emit16(d, p, d.master[p]);
}

// He-uniform weights (bound sqrt(6/fan_in)) from a counter hash, zero pads
// and biases, zero Adam moments
__global__ void k_nrc_init(NrcDevice d, uint64_t seed) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNrcParams) return;
  float w = 0.f;
  if (p < kNrcWeights) {
    int l = 5;
    while (l > 0 && p < nrc_woff(l)) --l;
    const int local = p - nrc_woff(l);
    const int o = local / nrc_in(l), i = local % nrc_in(l);
    const int fan_in = l == 0 ? kNrcInputs : kNrcWidth;
    const bool real = i < fan_in && (l < 5 ? o < kNrcWidth : o < 3);
    if (real) {
      uint64_t z = seed + 0x9E3779B97F4A7C15ull * uint64_t(p + 1);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      const float u = float(z >> 40) * (1.f / 16777216.f);
      w = (2.f * u - 1.f) * sqrtf(6.f / float(fan_in));
    }
  }
  d.master[p] = w;
  d.adam_m[p] = 0.f;
  d.adam_v[p] = 0.f;
  emit16(d, p, w);
  if (p < 4) d.stats[p] = 0.f;  // loss, samples, Adam step count
}

// one-wave MFMA lane-layout probe of the 16x16x32 instruction: C (16x16)
// = A (16x32) B (32x16) with the operand layout the kernels assume
template <int BF16>
__global__ void k_mfma_probe(const float *a, const float *b, float *c) {
  const int l = threadIdx.x & 63, r16 = l & 15, g = l >> 4;
  const float *ar = a + r16 * 32 + 8 * g;
  const Frag8 fa = cat(pack4(ar[0], ar[1], ar[2], ar[3], BF16),
                       pack4(ar[4], ar[5], ar[6], ar[7], BF16));
  float bv[8];
  for (int j = 0; j < 8; ++j) bv[j] = b[(8 * g + j) * 16 + r16];
  const Frag8 fb = cat(pack4(bv[0], bv[1], bv[2], bv[3], BF16),
                       pack4(bv[4], bv[5], bv[6], bv[7], BF16));
  f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
  acc = mfma<BF16>(fa, fb, acc);
  for (int j = 0; j < 4; ++j) c[(4 * g + j) * 16 + r16] = acc[j];
}

// ---------------------------------------------------------------------------
hipError_t nrc_launch_init(const NrcDevice &d, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_nrc_init, dim3((kNrcParams + 255) / 256), dim3(256), 0, s, d, seed);
  return hipGetLastError();
}

hipError_t nrc_launch_refresh(const NrcDevice &d, hipStream_t s) {
  hipLaunchKernelGGL(k_nrc_refresh, dim3((kNrcParams + 255) / 256), dim3(256), 0, s, d);
  return hipGetLastError();
}

hipError_t nrc_launch_infer(const NrcDevice &d, int n, const float *raw, float *out, int num_cus,
                            hipStream_t s, const uint32_t *n_dev) {
  if (n <= 0) return hipSuccess;
  const int tiles = (n + 16 * kInferQB - 1) / (16 * kInferQB);
  const int blocks = std::max(1, std::min((tiles + 3) / 4, 2 * std::max(1, num_cus)));
  if (d.bf16)
    hipLaunchKernelGGL(k_nrc_infer<1>, dim3(blocks), dim3(256), 0, s, d, n, n_dev, raw, out);
  else
    hipLaunchKernelGGL(k_nrc_infer<0>, dim3(blocks), dim3(256), 0, s, d, n, n_dev, raw, out);
  return hipGetLastError();
}

// The gradient pass of mini-batch `batch`, then the fixed-order reduction
// (ADAM: fused with the Adam step).  The workgroup count comes from the
// largest possible mini-batch; every workgroup writes its partial (zeros
// when it got no chunk).
template <int ADAM>
static hipError_t launch_grad_reduce(const NrcDevice &d, int n, const float *raw,
                                     const float *target, int max_workgroups, float lr, float b1,
                                     float b2, float eps, hipStream_t s, const uint32_t *n_dev,
                                     int batch, int nbatches) {
  const int chunks = (n / std::max(1, nbatches) + 1 + 63) / 64;
  const int parts = std::max(1, std::min(chunks, max_workgroups));
  if (d.bf16)
    hipLaunchKernelGGL(k_nrc_grad<1>, dim3(parts), dim3(256), 0, s, d, n, n_dev, batch, nbatches,
                       ADAM, raw, target);
  else
    hipLaunchKernelGGL(k_nrc_grad<0>, dim3(parts), dim3(256), 0, s, d, n, n_dev, batch, nbatches,
                       ADAM, raw, target);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_nrc_reduce<ADAM>, dim3((kNrcParams + 1 + 63) / 64), dim3(256), 0, s, d,
                     parts, n, n_dev, batch, nbatches, lr, b1, b2, eps);
  return hipGetLastError();
}

hipError_t nrc_launch_gradients(const NrcDevice &d, int n, const float *raw, const float *target,
                                int max_workgroups, hipStream_t s, const uint32_t *n_dev,
                                int batch, int nbatches) {
  return launch_grad_reduce<0>(d, n, raw, target, max_workgroups, 0.f, 0.f, 0.f, 0.f, s, n_dev,
                               batch, nbatches);
}

hipError_t nrc_launch_train_step(const NrcDevice &d, int n, const float *raw, const float *target,
                                 int max_workgroups, float lr, float b1, float b2, float eps,
                                 hipStream_t s, const uint32_t *n_dev, int batch, int nbatches) {
  return launch_grad_reduce<1>(d, n, raw, target, max_workgroups, lr, b1, b2, eps, s, n_dev,
                               batch, nbatches);
}

hipError_t nrc_launch_mfma_probe(const float *a, const float *b, float *c, int bf16,
                                 hipStream_t s) {
  if (bf16)
    hipLaunchKernelGGL(k_mfma_probe<1>, dim3(1), dim3(64), 0, s, a, b, c);
  else
    hipLaunchKernelGGL(k_mfma_probe<0>, dim3(1), dim3(64), 0, s, a, b, c);
  return hipGetLastError();
}

}  // namespace vrt
