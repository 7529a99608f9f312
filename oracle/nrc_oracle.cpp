/*
 * ORACLE -- test infrastructure, not product code.
 *
 * FP64 restatement of the neural radiance cache MLP (no counterpart in the
 * reference).  The network, encoding and loss are defined in include/vrt.h.
 * `precision` selects what the GPU rounds: 0 = fp16 and 1 = bf16 round the
 * encoded inputs, weights, hidden activations and back-propagated dZ to 16
 * bits (as the matrix-core operands are) and accumulate exactly; 2 = no
 * rounding at all (plain FP64 network, used for finite-difference checks).
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kWeights = 8192 + 4 * 4096 + 1024;
constexpr int kBiases = 5 * 64 + 16;
constexpr int kParams = kWeights + kBiases;
inline int in_dim(int l) { return l == 0 ? 128 : 64; }
inline int out_dim(int l) { return l == 5 ? 16 : 64; }
inline int woff(int l) { return l == 0 ? 0 : 8192 + (l - 1) * 4096; }
inline int boff(int l) { return kWeights + l * 64; }

float f32(double x) { return float(x); }

// float -> IEEE half, round to nearest even, back to float
float round_half(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t sign = u & 0x80000000u;
  uint32_t a = u & 0x7FFFFFFFu;
  if (a >= 0x7F800000u) return f;            // inf / nan
  if (a >= 0x477FF000u) {                    // rounds to >= 65520: overflow
    uint32_t r = sign | 0x7F800000u;
    float o;
    std::memcpy(&o, &r, 4);
    return o;
  }
  float af;
  std::memcpy(&af, &a, 4);
  if (af < 6.103515625e-05f) {  // subnormal half: quantum 2^-24
    const double q = std::ldexp(1.0, -24);
    double v = std::nearbyint(double(af) / q) * q;  // ties-to-even (default mode)
    float o = float(v);
    return sign ? -o : o;
  }
  // normal: keep 10 mantissa bits
  const uint32_t lsb = (a >> 13) & 1u;
  a += 0xFFFu + lsb;
  a &= ~0x1FFFu;
  uint32_t r = sign | a;
  float o;
  std::memcpy(&o, &r, 4);
  return o;
}

float round_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return f;
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float o;
  std::memcpy(&o, &u, 4);
  return o;
}

double q16(double x, int precision) {
  if (precision == 0) return round_half(f32(x));
  if (precision == 1) return round_bf16(f32(x));
  return x;
}

// sin/cos(2^k pi x) = sin/cos(2 pi frac(x 2^(k-1))), the fraction formed in
// float exactly as the GPU does
void encode(const float *raw, double *e) {
  for (int f = 0; f < 128; ++f) {
    if (f < 120) {
      const int s = f / 24, w = f % 24, k = w >> 1;
      float r = raw[s] * std::ldexp(1.0f, k - 1);
      r = r - std::floor(r);
      const double ang = 2.0 * M_PI * double(r);
      e[f] = (w & 1) ? std::cos(ang) : std::sin(ang);
    } else {
      e[f] = f == 120 ? double(raw[5]) : 0.0;
    }
  }
}

struct Net {
  const float *p;
  int prec;
  double W(int l, int o, int i) const { return q16(p[woff(l) + o * in_dim(l) + i], prec); }
  double B(int l, int o) const { return double(p[boff(l) + o]); }
};

// forward of one sample; acts[l] = input of layer l (after rounding)
void forward(const Net &net, const float *raw, std::vector<double> acts[6], double out[16]) {
  double e[128];
  encode(raw, e);
  acts[0].assign(128, 0.0);
  for (int i = 0; i < 128; ++i) acts[0][i] = q16(e[i], net.prec);
  for (int l = 0; l < 6; ++l) {
    std::vector<double> z(out_dim(l));
    for (int o = 0; o < out_dim(l); ++o) {
      double s = net.B(l, o);
      for (int i = 0; i < in_dim(l); ++i) s += net.W(l, o, i) * acts[l][i];
      z[o] = s;
    }
    if (l < 5) {
      acts[l + 1].assign(64, 0.0);
      // fp32 accumulator -> ReLU -> 16-bit operand of the next layer
      for (int o = 0; o < 64; ++o) {
        const double zz = net.prec == 2 ? z[o] : double(f32(z[o]));
        acts[l + 1][o] = q16(std::max(0.0, zz), net.prec);
      }
    } else {
      for (int o = 0; o < 16; ++o) out[o] = z[o];
    }
  }
}

}  // namespace

extern "C" {

int orc_nrc_param_count(void) { return kParams; }

void orc_nrc_encode(const float *raw, double *out) { encode(raw, out); }

void orc_nrc_forward(const float *params, int precision, int n, const float *raw, double *out3) {
  Net net{params, precision};
  std::vector<double> acts[6];
  double out[16];
  for (int q = 0; q < n; ++q) {
    forward(net, raw + 6 * q, acts, out);
    for (int c = 0; c < 3; ++c) out3[3 * q + c] = out[c];
  }
}

// Gradient sums of L = sum_q sum_c (p-t)^2 / (p^2 + 0.01), p^2 held constant.
// Returns the loss sum.  dZ is rounded to 16 bits where the GPU rounds it
// (the staged dZ^T operands), masks come from the rounded activations.
double orc_nrc_gradients(const float *params, int precision, int n, const float *raw,
                         const float *target, double *grad) {
  Net net{params, precision};
  for (int i = 0; i < kParams; ++i) grad[i] = 0.0;
  double loss = 0.0;
  std::vector<double> acts[6];
  double out[16];
  for (int q = 0; q < n; ++q) {
    forward(net, raw + 6 * q, acts, out);
    std::vector<double> dz(16, 0.0);
    for (int c = 0; c < 3; ++c) {
      const double p = precision == 2 ? out[c] : double(f32(out[c])), t = target[3 * q + c];
      const double den = p * p + 0.01;
      loss += (p - t) * (p - t) / den;
      dz[c] = 2.0 * (p - t) / den;
    }
    for (int l = 5; l >= 0; --l) {
      std::vector<double> dzq(out_dim(l));
      for (int o = 0; o < out_dim(l); ++o) dzq[o] = q16(dz[o], precision);
      for (int o = 0; o < out_dim(l); ++o) {
        grad[boff(l) + o] += dz[o];
        for (int i = 0; i < in_dim(l); ++i) grad[woff(l) + o * in_dim(l) + i] += dzq[o] * acts[l][i];
      }
      if (l > 0) {
        std::vector<double> dx(in_dim(l), 0.0);
        for (int i = 0; i < in_dim(l); ++i) {
          double s = 0.0;
          for (int o = 0; o < out_dim(l); ++o) s += net.W(l, o, i) * dzq[o];
          dx[i] = acts[l][i] > 0.0 ? s : 0.0;
        }
        dz = dx;
      }
    }
  }
  return loss;
}

}  // extern "C"
