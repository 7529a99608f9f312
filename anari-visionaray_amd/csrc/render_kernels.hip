// gfx950 kernels of the per-pixel rendering path: camera ray generation,
// brick-ordered trilinear fetch, TF classification from LDS, emission-
// absorption ray marching (raycast), majorant-grid DDA + Woodcock delta
// tracking and the path-tracer state machine (pathtrace), and frame
// accumulation / tone mapping.
//
// Launch shape: 256-thread workgroups cover 16x16 pixel tiles; each 64-lane
// wave owns an 8x8 sub-tile so the rays of a wave stay spatially coherent
// (the reference launches 64x1 rows, for_each.h:272-280).  Workgroup ids are
// remapped so each of the 8 XCDs (one L2 each) renders one contiguous band of
// tiles and re-reads the same volume bricks from its own L2.
#include "render_kernels.h"

#include <float.h>

#ifndef VRT_NATIVE_TRIG
#define VRT_NATIVE_TRIG 0
#endif
#ifndef VRT_FAST_RCP
#define VRT_FAST_RCP 0
#endif

namespace vrt {

// reciprocal of a ray direction component (slab tests, DDA setup)
__device__ inline float rcp_dir(float x) {
#if VRT_FAST_RCP
  return x == 0.f ? copysignf(__builtin_inff(), x) : __builtin_amdgcn_rcpf(x);
#else
  return 1.f / x;
#endif
}

// ---------------------------------------------------------------------------
// RNG: LCG<4> seeded by TEA-4 (common.h:55-102)
// ---------------------------------------------------------------------------
__device__ inline uint32_t tea4(uint32_t v0, uint32_t v1) {
  uint32_t s0 = 0;
#pragma unroll
  for (int n = 0; n < 4; n++) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
__device__ inline float rnd(uint32_t &s) {
  s = 1664525u * s + 1013904223u;
  return float(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// Work queues: the persistent kernels cut the image into 8 parts, one per
// XCD (the hardware places workgroup b on XCD b % 8), each drained from an
// atomic counter
// ---------------------------------------------------------------------------
constexpr int kQueues = 8;

// ---------------------------------------------------------------------------
// Rays / boxes
// ---------------------------------------------------------------------------
struct KRay {
  f3 ori, dir;
  float tmin, tmax;
};

struct BoxHit {
  bool hit;
  float tnear, tfar;
};

__device__ inline BoxHit intersect_box(const KRay &r, f3 lo, f3 hi) {
  f3 inv = mk(rcp_dir(r.dir.x), rcp_dir(r.dir.y), rcp_dir(r.dir.z));
  f3 t1 = (lo - r.ori) * inv;
  f3 t2 = (hi - r.ori) * inv;
  float tn = fmaxf(fminf(t1.z, t2.z), fmaxf(fminf(t1.y, t2.y), fminf(t1.x, t2.x)));
  float tf = fminf(fmaxf(t1.z, t2.z), fminf(fmaxf(t1.y, t2.y), fmaxf(t1.x, t2.x)));
  return BoxHit{tf >= tn && tf >= 0.f, tn, tf};
}

// renderer/common.h:14-18
__device__ inline float epsilon_from(f3 P, f3 dir, float t) {
  return fmaxf(max3(fabs3(P)), max3(fabs3(dir)) * t) * 0x1.fp-18f;
}

__device__ inline f3 mulA(const float *A, f3 p) {
  return mk(A[0] * p.x + A[1] * p.y + A[2] * p.z, A[3] * p.x + A[4] * p.y + A[5] * p.z,
            A[6] * p.x + A[7] * p.y + A[8] * p.z);
}

__device__ inline KRay to_local(const KVolume &v, const KRay &r) {
  KRay l = r;
  l.ori = mulA(v.A, r.ori + v.tinv);
  l.dir = mulA(v.A, r.dir);
  return l;
}

// ---------------------------------------------------------------------------
// Field fetch: trilinear, texel-centre convention (p*N-0.5), clamp
// addressing, on the brick records of vrt_structured_bricks: the cell index
// and weight per axis give one brick record holding all 8 taps at fixed
// offsets from one address.
// ---------------------------------------------------------------------------
// Clamping c to [0, n-1] gives the clamp-addressed taps exactly: below 0
// both taps are voxel 0 (the weight 0 keeps it), above n-1 both are voxel
// n-1 (the record repeats it); NaN clamps to 0 as the tap clamp did.
__device__ inline int axis_cell(float p, int n, float &w) {
  float c = p * float(n) - 0.5f;
  c = fminf(fmaxf(c, 0.f), float(n - 1));
  const float fl = floorf(c);
  w = c - fl;
  return int(fl);
}

__device__ inline float lerpf(float a, float b, float t) { return a + t * (b - a); }
// a + t (b - a) with the product and sum fused (one rounding): the structured
// field's trilinear filter
__device__ inline float lerp_fma(float a, float b, float t) { return __builtin_fmaf(t, b - a, a); }

// x^y from IEEE basic operations (fused multiply-add, ldexp, rint) and a
// table, so the opacity correction powf(1-alpha, dt/unitDistance) of
// VolumeIntegration.h:416-417 rounds identically on every machine that
// evaluates it this way.  log2: x = 2^e z with z in [0.69921875,
// 1.3984375) (float bits offset by 0x3f330000), z's subinterval i (the next
// 6 mantissa bits) gives {1/c_i as float, -log2 of that float}, r = z/c_i - 1
// = fma(z, 1/c_i, -1) with |r| <= 2^-7, log2(1 + r) from its Taylor series
// to r^4; c_i is the subinterval's midpoint, or 1 exactly for the one holding
// 1.0, so x near 1 keeps its digits.  2^(y log2 x) = 2^n 2^f with n =
// rint, |f| <= 1/2, 2^f from its Taylor series to f^7, scaled by ldexp.
// Error <= 1.5 (2 + |y log2 x|) ulp for normal x, y finite; results below
// 2^-126 are returned as 0.
__constant__ float2 kPowLog2[64] = {
{1.4222222566604614f, -0.5081469416618347f},
    {1.4065934419631958f, -0.4922053813934326f},
    {1.39130437374115f, -0.47643807530403137f},
    {1.376344084739685f, -0.46084117889404297f},
    {1.3617020845413208f, -0.4454111158847809f},
    {1.3473684787750244f, -0.43014445900917053f},
    {1.3333333730697632f, -0.4150375425815582f},
    {1.3195875883102417f, -0.4000871181488037f},
    {1.3061224222183228f, -0.38529011607170105f},
    {1.2929292917251587f, -0.37064337730407715f},
    {1.2799999713897705f, -0.3561437726020813f},
    {1.2673267126083374f, -0.34178850054740906f},
    {1.2549020051956177f, -0.3275747001171112f},
    {1.2427184581756592f, -0.31349948048591614f},
    {1.2307692766189575f, -0.2995603382587433f},
    {1.2190476655960083f, -0.285754531621933f},
    {1.2075471878051758f, -0.27207955718040466f},
    {1.1962616443634033f, -0.25853297114372253f},
    {1.185185194015503f, -0.24511250853538513f},
    {1.174311876296997f, -0.23181560635566711f},
    {1.163636326789856f, -0.21864023804664612f},
    {1.1531531810760498f, -0.20558416843414307f},
    {1.1428571939468384f, -0.1926451474428177f},
    {1.1327433586120605f, -0.17982102930545807f},
    {1.1228070259094238f, -0.16710999608039856f},
    {1.1130434274673462f, -0.15450988709926605f},
    {1.1034482717514038f, -0.14201900362968445f},
    {1.0940171480178833f, -0.12963534891605377f},
    {1.0847457647323608f, -0.11735695600509644f},
    {1.0756303071975708f, -0.10518231242895126f},
    {1.0666667222976685f, -0.09310948103666306f},
    {1.0578511953353882f, -0.08113670349121094f},
    {1.049180269241333f, -0.06926257908344269f},
    {1.0406503677368164f, -0.057485442608594894f},
    {1.0322580337524414f, -0.04580364748835564f},
    {1.0240000486373901f, -0.034215785562992096f},
    {1.0158730745315552f, -0.022720159962773323f},
    {1.0078740119934082f, -0.011315307579934597f},
    {1.0f, -0.0f},
    {0.9846153855323792f, 0.0223678108304739f},
    {0.9696969985961914f, 0.044394075870513916f},
    {0.9552238583564758f, 0.0660892203450203f},
    {0.9411764740943909f, 0.08746283501386642f},
    {0.9275362491607666f, 0.10852442681789398f},
    {0.9142857193946838f, 0.12928301095962524f},
    {0.9014084339141846f, 0.14974714815616608f},
    {0.8888888955116272f, 0.16992498934268951f},
    {0.8767123222351074f, 0.18982456624507904f},
    {0.8648648858070374f, 0.20945332944393158f},
    {0.8533333539962769f, 0.22881865501403809f},
    {0.8421052694320679f, 0.2479275017976761f},
    {0.8311688303947449f, 0.2667865455150604f},
    {0.8205128312110901f, 0.28540220856666565f},
    {0.8101266026496887f, 0.3037807047367096f},
    {0.800000011920929f, 0.32192808389663696f},
    {0.790123462677002f, 0.33984997868537903f},
    {0.7804877758026123f, 0.35755205154418945f},
    {0.7710843086242676f, 0.3750394880771637f},
    {0.761904776096344f, 0.39231738448143005f},
    {0.7529411911964417f, 0.4093908965587616f},
    {0.7441860437393188f, 0.42626476287841797f},
    {0.7356321811676025f, 0.4429435133934021f},
    {0.7272727489471436f, 0.45943158864974976f},
    {0.7191011309623718f, 0.4757334291934967f}};

__device__ inline float pow_ieee(float x, float y, const float2 *tab) {
  if (y == 0.f || x == 1.f) return 1.f;
  if (!(x > 0.f)) return y > 0.f ? 0.f : __builtin_inff();
  const uint32_t bits = __float_as_uint(x);
  if (bits < 0x00800000u) return y > 0.f ? 0.f : __builtin_inff();  // subnormal base
  if (bits >= 0x7f800000u) return x;                                 // inf
  // the exponents of the usual sampling rates (2 and 0.5 samples per unit
  // distance) take one correctly rounded operation
  if (y == 0.5f) return sqrtf(x);
  if (y == 2.f) {
    const float s = x * x;
    return s < 1.17549435e-38f ? 0.f : s;
  }
  const uint32_t tmp = bits - 0x3f330000u;
  const int e = int(tmp) >> 23;
  const float2 t = tab[(tmp >> 17) & 63u];
  const float z = __uint_as_float(bits - (tmp & 0xff800000u));
  const float r = __builtin_fmaf(z, t.x, -1.f);
  float q = __builtin_fmaf(r, -0.36067376f, 0.48089835f);  // -1/(4 ln2), 1/(3 ln2)
  q = __builtin_fmaf(r, q, -0.72134752f);                   // -1/(2 ln2)
  q = __builtin_fmaf(r, q, 1.44269504f);                    // 1/ln2
  const float l2 = __builtin_fmaf(r, q, float(e) + t.y);
  const float zz = y * l2;
  if (zz < -126.f) return 0.f;
  if (zz > 128.f) return __builtin_inff();
  const float n = __builtin_rintf(zz);
  const float f = zz - n;
  float p = 1.52527338e-5f;                  // ln2^7 / 7!
  p = __builtin_fmaf(p, f, 1.54035304e-4f);  // ln2^6 / 6!
  p = __builtin_fmaf(p, f, 1.33335581e-3f);
  p = __builtin_fmaf(p, f, 9.61812911e-3f);
  p = __builtin_fmaf(p, f, 5.55041087e-2f);
  p = __builtin_fmaf(p, f, 0.240226507f);
  p = __builtin_fmaf(p, f, 0.693147181f);
  p = __builtin_fmaf(p, f, 1.f);
  return __builtin_ldexpf(p, int(n));
}

__device__ inline float sample_structured(const KVolume &v, f3 P) {
  float wx, wy, wz;
  const int i = axis_cell(P.x, v.dims[0], wx);
  const int j = axis_cell(P.y, v.dims[1], wy);
  const int k = axis_cell(P.z, v.dims[2], wz);
  const uint32_t brick = (uint32_t(k >> 3) * v.nby + uint32_t(j >> 3)) * v.nbx + uint32_t(i >> 3);
  const float *b = v.bricks + size_t(brick) * VRT_BRICK_FLOATS +
                   uint32_t(((k & 7) * 9 + (j & 7)) * 9 + (i & 7));
  const float v000 = b[0], v100 = b[1], v010 = b[9], v110 = b[10];
  const float v001 = b[81], v101 = b[82], v011 = b[90], v111 = b[91];
  float c00 = lerp_fma(v000, v100, wx), c10 = lerp_fma(v010, v110, wx);
  float c01 = lerp_fma(v001, v101, wx), c11 = lerp_fma(v011, v111, wx);
  return lerp_fma(lerp_fma(c00, c10, wy), lerp_fma(c01, c11, wy), wz);
}

// ---------------------------------------------------------------------------
// NanoVDB field fetch (sampleField, DeviceCopyableObjects.h:554-569) on the
// two-level layout of nanovdb_field.cpp: NanoGrid<float>::getValue is a top
// read (128^3 region: block or constant), a block read (8^3 slot: leaf or
// constant) and, in a leaf, one value read.
// ---------------------------------------------------------------------------
// entry of the 8^3 slot (sx, sy, sz) (voxel index >> 3): >= 0 leaf index,
// else -1 - constant index
__device__ inline int nvdb_slot(const KVolume &v, int sx, int sy, int sz) {
  const int tx = (sx >> 4) - v.nv_lo[0], ty = (sy >> 4) - v.nv_lo[1], tz = (sz >> 4) - v.nv_lo[2];
  if (uint32_t(tx) >= uint32_t(v.nv_n[0]) || uint32_t(ty) >= uint32_t(v.nv_n[1]) ||
      uint32_t(tz) >= uint32_t(v.nv_n[2]))
    return -1;  // background
  const int t = v.nv_top[(size_t(tz) * size_t(v.nv_n[1]) + size_t(ty)) * size_t(v.nv_n[0]) + tx];
  if (t < 0) return t;
  return v.nv_blocks[size_t(t) * 4096u + uint32_t((((sz & 15) << 4) | (sy & 15)) << 4 | (sx & 15))];
}

__device__ inline float nvdb_value(const KVolume &v, int i, int j, int k) {
  const int e = nvdb_slot(v, i >> 3, j >> 3, k >> 3);
  return e >= 0 ? v.leaves[size_t(e) * 512u + uint32_t(((i & 7) << 6) | ((j & 7) << 3) | (k & 7))]
                : v.nv_consts[-1 - e];
}

// createSampler<1|0>(acc)(worldToIndexF(P)): the map is an fmaf chain over
// P - translation (Map::applyInverseMapF); the trilinear sampler lerps over
// z, then y, then x with lerp(a, b, w) = a + w (b - a) and weights
// idx - floor(idx) (math/SampleFromVoxels.h:174-290, 324-430); the nearest
// sampler reads the voxel at rint(idx)
__device__ inline float sample_nanovdb(const KVolume &v, f3 P) {
  const float x = P.x - v.nv_vec[0], y = P.y - v.nv_vec[1], z = P.z - v.nv_vec[2];
  float idx[3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
    idx[a] = __builtin_fmaf(x, v.nv_inv[3 * a],
                            __builtin_fmaf(y, v.nv_inv[3 * a + 1], z * v.nv_inv[3 * a + 2]));
  if (v.filter == VRT_FILTER_NEAREST)
    return nvdb_value(v, int(__builtin_rintf(idx[0])), int(__builtin_rintf(idx[1])),
                      int(__builtin_rintf(idx[2])));
  const float fx = floorf(idx[0]), fy = floorf(idx[1]), fz = floorf(idx[2]);
  const float wx = idx[0] - fx, wy = idx[1] - fy, wz = idx[2] - fz;
  const int i = int(fx), j = int(fy), k = int(fz);
  // the accessor's leaf cache: the slot of the (i, j, k) corner serves every
  // tap in the same 8^3 slot; a tap across a slot boundary looks its own up
  const int e0 = nvdb_slot(v, i >> 3, j >> 3, k >> 3);
  auto tap = [&](int a, int b, int c) {
    int e = e0;
    if (((a >> 3) != (i >> 3)) | ((b >> 3) != (j >> 3)) | ((c >> 3) != (k >> 3)))
      e = nvdb_slot(v, a >> 3, b >> 3, c >> 3);
    return e >= 0 ? v.leaves[size_t(e) * 512u + uint32_t(((a & 7) << 6) | ((b & 7) << 3) | (c & 7))]
                  : v.nv_consts[-1 - e];
  };
  const float v00 = lerpf(tap(i, j, k), tap(i, j, k + 1), wz);
  const float v01 = lerpf(tap(i, j + 1, k), tap(i, j + 1, k + 1), wz);
  const float v10 = lerpf(tap(i + 1, j, k), tap(i + 1, j, k + 1), wz);
  const float v11 = lerpf(tap(i + 1, j + 1, k), tap(i + 1, j + 1, k + 1), wz);
  return lerpf(lerpf(v00, v01, wy), lerpf(v10, v11, wy), wx);
}

// ---------------------------------------------------------------------------
// Unstructured field fetch (sampleField, DeviceCopyableObjects.h:482-527).
// ---------------------------------------------------------------------------
// intersectTet (Plane.h:56-96) with its arithmetic and test order: vertices
// relative to P, the plane through three of them (N = cross(b-a, c-a),
// d = dot(a, N)), the ratio eval(0)/eval(opposite vertex) as the barycentric
// weight; each weight must be >= 0 and the first <= 1.  With no data at the
// first vertex the value is the cell value, kept in the second vertex's w
// (DeviceCopyableObjects.h:150-153).
__device__ inline float plane_ratio(f3 a, f3 b, f3 c, f3 opp) {
  const f3 n = cross(b - a, c - a);
  const float d = dot(a, n);
  return (dot(mk(0.f), n) - d) / (dot(opp, n) - d);
}

__device__ inline bool tet_value(f3 P, float4 A, float4 B, float4 C, float4 D, float &value) {
  const f3 va = mk(A.x, A.y, A.z) - P, vb = mk(B.x, B.y, B.z) - P;
  const f3 vc = mk(C.x, C.y, C.z) - P, vd = mk(D.x, D.y, D.z) - P;
  const float fa = plane_ratio(vb, vd, vc, va);
  if (fa < 0.f || fa > 1.f) return false;
  const float fb = plane_ratio(va, vc, vd, vb);
  if (fb < 0.f) return false;
  const float fc = plane_ratio(va, vd, vb, vc);
  if (fc < 0.f) return false;
  const float fd = plane_ratio(va, vb, vc, vd);
  if (fd < 0.f) return false;
  value = __builtin_isnan(A.w) ? B.w : fa * A.w + fb * B.w + fc * C.w + fd * D.w;
  return true;
}

// Points inside the field bounds, faces included (the element BVH's root box
// test is inclusive), map to the grid cell of P (float arithmetic the list
// construction allows for, the top face clamped into the last cell); its
// tetrahedra are tested in list order until one contains P.
__device__ inline bool sample_tets(const KVolume &v, f3 P, float &value, uint32_t &rec) {
  if (!(P.x >= v.blo.x && P.x <= v.bhi.x && P.y >= v.blo.y && P.y <= v.bhi.y &&
        P.z >= v.blo.z && P.z <= v.bhi.z))
    return false;  // also on NaN
  // the grid origin is the bounds' lower corner, so q >= 0
  const uint32_t ix = min(uint32_t((P.x - v.ug_lo.x) * v.ug_inv.x), uint32_t(v.ug_dims[0] - 1));
  const uint32_t iy = min(uint32_t((P.y - v.ug_lo.y) * v.ug_inv.y), uint32_t(v.ug_dims[1] - 1));
  const uint32_t iz = min(uint32_t((P.z - v.ug_lo.z) * v.ug_inv.z), uint32_t(v.ug_dims[2] - 1));
  const uint32_t c = (iz * uint32_t(v.ug_dims[1]) + iy) * uint32_t(v.ug_dims[0]) + ix;
  const uint32_t b = v.ug_start[c], e = v.ug_start[c + 1];
  for (uint32_t i = b; i < e; ++i) {
    const float4 *r = v.ug_recs + size_t(i) * 4u;
    if (tet_value(P, r[0], r[1], r[2], r[3], value)) {
      rec = i;
      return true;
    }
  }
  return false;
}

// sampleField: voxel-space point -> field value; false where the field has
// no value (outside every cell of an unstructured field).  ANYF = false
// compiles the structuredRegular fetch alone (scenes without other fields).
// `rec`: the list entry of the containing tetrahedron (kNoRec for the other
// fields, whose primitive id is 0, DeviceCopyableObjects.h:478-481, 556-567).
constexpr uint32_t kNoRec = 0xFFFFFFFFu;

template <bool ANYF = true>
__device__ inline bool sample_field(const KVolume &v, f3 P, float &value, uint32_t &rec) {
  rec = kNoRec;
  if (ANYF && v.field_type == VRT_FIELD_UNSTRUCTURED) return sample_tets(v, P, value, rec);
  value = (ANYF && v.field_type == VRT_FIELD_NANOVDB) ? sample_nanovdb(v, P)
                                                      : sample_structured(v, P);
  return true;
}

template <bool ANYF = true>
__device__ inline bool sample_field(const KVolume &v, f3 P, float &value) {
  uint32_t rec;
  return sample_field<ANYF>(v, P, value, rec);
}

// sampleGradient (DeviceCopyableObjects.h:582-597): central differences at
// P +- delta along each axis, false unless all six samples have a value.
// The gradient normal is its normalisation; a zero gradient (a flat region)
// gives no normal here -- the reference's normalize(0) is NaN.
template <bool ANYF = true>
__device__ inline f3 gradient_normal(const KVolume &v, f3 P) {
  const f3 delta = v.S * v.cell_size;  // cellSize * diag(voxelSpaceTransform)
  float x0, x1, y0, y1, z0, z1;
  const bool b0 = sample_field<ANYF>(v, P + mk(delta.x, 0.f, 0.f), x1);
  const bool b1 = sample_field<ANYF>(v, P - mk(delta.x, 0.f, 0.f), x0);
  const bool b2 = sample_field<ANYF>(v, P + mk(0.f, delta.y, 0.f), y1);
  const bool b3 = sample_field<ANYF>(v, P - mk(0.f, delta.y, 0.f), y0);
  const bool b4 = sample_field<ANYF>(v, P + mk(0.f, 0.f, delta.z), z1);
  const bool b5 = sample_field<ANYF>(v, P - mk(0.f, 0.f, delta.z), z0);
  if (!(b0 && b1 && b2 && b3 && b4 && b5)) return mk(0.f);
  const f3 g = mk(x1, y1, z1) - mk(x0, y0, z0);
  if (g.x == 0.f && g.y == 0.f && g.z == 0.f) return mk(0.f);
  return normalize(g);
}

// postClassify (DeviceCopyableObjects.h:615-619): tex1D(linear, clamp) of the
// TF table, read from the workgroup's LDS copy
// the two table entries and the weight tex1D(linear, clamp) reads for val
__device__ inline void tf_taps(const KVolume &v, float val, int &i0, int &i1, float &w) {
  // a power-of-two range divides exactly as its reciprocal multiplies
  float u = v.vrange_pow2 ? (val - v.vlo) * v.vrange_inv : (val - v.vlo) / (v.vhi - v.vlo);
  const int n = v.tf_size;
  float c = u * float(n) - 0.5f;
  c = fminf(fmaxf(c, -1.f), float(n));
  float fl = floorf(c);
  w = c - fl;
  int i = int(fl);
  i0 = v.tf_lds_offset + min(max(i, 0), n - 1);
  i1 = v.tf_lds_offset + min(max(i + 1, 0), n - 1);
}
__device__ inline float4 classify(const KVolume &v, const float4 *lds, float val) {
  int i0, i1;
  float w;
  tf_taps(v, val, i0, i1, w);
  const float4 a = lds[i0], b = lds[i1];
  return make_float4(lerpf(a.x, b.x, w), lerpf(a.y, b.y, w), lerpf(a.z, b.z, w),
                     lerpf(a.w, b.w, w));
}
// opacity alone (the component delta tracking tests), as classify computes it
__device__ inline float classify_alpha(const KVolume &v, const float4 *lds, float val) {
  int i0, i1;
  float w;
  tf_taps(v, val, i0, i1, w);
  const float *t = reinterpret_cast<const float *>(lds);
  return lerpf(t[4 * i0 + 3], t[4 * i1 + 3], w);
}

// Workgroup LDS: [TF tables, float4 each][majorant steps, one float4][majorant
// bytes of the grids that fit]
__host__ __device__ inline size_t scene_lds_bytes(const KScene &sc) {
  return size_t(sc.tf_lds_total) * 16 + 16 + ((size_t(sc.maj_lds_bytes) + 15) & ~size_t(15));
}
__device__ inline float *lds_steps(const KScene &sc, float4 *lds) {
  return reinterpret_cast<float *>(lds + sc.tf_lds_total);
}
__device__ inline const uint8_t *lds_majorants(const KScene &sc, const float4 *lds, int vi) {
  const KVolume &v = sc.vol[vi];
  if (v.maj_lds_offset < 0) return nullptr;
  return reinterpret_cast<const uint8_t *>(lds + sc.tf_lds_total + 1) + v.maj_lds_offset;
}

__device__ inline void stage_tf(const KScene &sc, float4 *lds) {
  float *steps = lds_steps(sc, lds);
  uint8_t *bytes = reinterpret_cast<uint8_t *>(lds + sc.tf_lds_total + 1);
  for (int vi = 0; vi < sc.num_volumes; ++vi) {
    const KVolume &v = sc.vol[vi];
    for (int i = threadIdx.x; i < v.tf_size; i += blockDim.x) lds[v.tf_lds_offset + i] = v.tf[i];
    if (v.maj_lds_offset >= 0) {
      const int n = v.gdims[0] * v.gdims[1] * v.gdims[2];
      uint8_t *dst = bytes + v.maj_lds_offset;
      const int n16 = n >> 4;
      for (int i = threadIdx.x; i < n16; i += blockDim.x)
        reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(v.mbytes)[i];
      for (int i = (n16 << 4) + threadIdx.x; i < n; i += blockDim.x) dst[i] = v.mbytes[i];
    }
    if (threadIdx.x == 0) steps[vi] = v.mstep ? *v.mstep : 0.f;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Camera (ANARI perspective; see vrt_camera)
// ---------------------------------------------------------------------------
__device__ inline KRay primary_ray(const KCamera &c, float x, float y, float W, float H) {
  float sx = (x + 0.5f) / W, sy = (y + 0.5f) / H;
  sx = (1.f - sx) * c.region[0] + sx * c.region[2];
  sy = (1.f - sy) * c.region[1] + sy * c.region[3];
  f3 d = c.W + c.U * (2.f * sx - 1.f) + c.V * (2.f * sy - 1.f);
  KRay r;
  r.ori = c.pos;
  r.dir = normalize(d);
  r.tmin = 0.f;
  r.tmax = FLT_MAX;
  return r;
}

// clipRay (renderer/common.h:1469-1482): every plane (N, D) keeps N.x <= D
__device__ inline void clip_ray(KRay &r, const KParams &P) {
  for (int i = 0; i < P.num_clip_planes; ++i) {
    const float4 c = P.clip[i];
    const f3 n = mk(c.x, c.y, c.z);
    const float s = dot(n, r.dir);
    if (s != 0.f) {
      const float t = (c.w - dot(n, r.ori)) / s;
      if (s < 0.f)
        r.tmin = fmaxf(r.tmin, t);
      else
        r.tmax = fminf(r.tmax, t);
    }
  }
}

// Per-lane picks from the scene's volumes, lights or surfaces (x =
// sc.vol[id].field for a lane-dependent id) are written as a loop over
// every entry whose loads do not depend on the comparison: x = (i == id) ?
// sc.vol[i].field : x, with i made opaque to the compiler.  Otherwise the
// compiler folds the loop back into sc.vol[id] -- a divergent index into the
// kernel argument, which copies the whole scene to scratch memory.
__device__ inline int uidx(int i) { return __builtin_amdgcn_readfirstlane(i); }
// the same index once a comparison with the lane's id has selected it: an
// empty asm hides the equality, so the compiler cannot address the entry
// through the lane's id instead
__device__ inline int opaque_uidx(int i) {
  asm volatile("" : "+s"(i));
  return i;
}

// ---------------------------------------------------------------------------
// Triangles and the 4-wide BVH (bvh_build.cpp), quad lights
// ---------------------------------------------------------------------------
// basic_triangle {v1, e1, e2} (visionaray, not vendored): Moller-Trumbore
// with inclusive edges; t is returned unchecked
__device__ inline bool tri_test(const KRay &r, f3 v1, f3 e1, f3 e2, float &t, float &u,
                                float &v) {
  const f3 s1 = cross(r.dir, e2);
  const float div = dot(s1, e1);
  if (div == 0.f) return false;
  const float inv = 1.f / div;
  const f3 d = r.ori - v1;
  const float b1 = dot(d, s1) * inv;
  if (b1 < 0.f || b1 > 1.f) return false;
  const f3 s2 = cross(d, e1);
  const float b2 = dot(r.dir, s2) * inv;
  if (b2 < 0.f || b1 + b2 > 1.f) return false;
  t = dot(e2, s2) * inv;
  u = b1;
  v = b2;
  return true;
}

constexpr int kBvhStack = VRT_BVH_MAX_STACK;  // entries per lane (builds check the bound)
constexpr int kBvhLdsStack = 16;  // of which the first (top of a shallow stack) live in LDS
constexpr int kBvhLanes = 256;    // lanes of the workgroups that traverse (LDS stride)

// LDS bytes of the traversal stacks of one 256-lane workgroup
__host__ __device__ constexpr size_t bvh_lds_stack_bytes() {
  return size_t(kBvhLdsStack) * kBvhLanes * sizeof(int2);
}

// BVH nodes staged in LDS: every ray enters at the root, so the top levels
// of each mesh's tree (the first nodes of its breadth-first array) are read
// by all lanes; a workgroup copies them once and reads them from LDS.
constexpr int kMaxStagedNodes = 256;  // 32 KB per workgroup at most
constexpr size_t kLdsBytesPerCu = 160 * 1024;
__host__ __device__ inline size_t staged_nodes_bytes(const KScene &sc) {
  size_t n = 0;
  for (int i = 0; i < sc.num_surfaces; ++i) n += size_t(sc.staged[i].lds_nodes);
  return n * 8 * sizeof(float4);
}
// the staged nodes follow the tables and the traversal stacks
__host__ __device__ inline size_t staged_nodes_offset(const KScene &sc) {
  return scene_lds_bytes(sc) + bvh_lds_stack_bytes();
}
__device__ inline const float4 *staged_nodes(const KScene &sc, const float4 *lds) {
  return reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(lds) +
                                          staged_nodes_offset(sc));
}
// copy of the staged nodes (the caller's barrier publishes it)
__device__ inline void stage_nodes(const KScene &sc, float4 *lds) {
  float4 *dst = const_cast<float4 *>(staged_nodes(sc, lds));
  for (int si = 0; si < sc.num_surfaces; ++si) {
    const float4 *src = sc.surf[si].nodes;
    const int first = sc.staged[si].lds_first;
    for (int i = threadIdx.x; i < 8 * sc.staged[si].lds_nodes; i += blockDim.x)
      dst[first + i] = src[i];
  }
}

// compare-exchange of (node, entry distance) pairs by distance
__device__ inline void cswap(float &ta, int &na, float &tb, int &nb) {
  if (tb < ta) {
    const float t = ta;
    ta = tb;
    tb = t;
    const int n = na;
    na = nb;
    nb = n;
  }
}

// Closest hit in [tmin, tbest] (tbest in: the far end), or any hit when
// `any`; ties go to the lower triangle index, so the result does not depend
// on the traversal order.  Boxes are tested with slab distances from finite
// reciprocals (a zero direction component becomes +-1e20), which can only
// add box hits, never drop one.  The traversal descends into the nearest
// child box that was hit and stacks the others (node, entry distance); the
// stack's first kBvhLdsStack entries live in LDS (`lst`: this lane's slot,
// entries kBvhLanes apart, so the lanes of a wave hit distinct banks), the
// rest in private memory, which only deep traversals reach.  Nodes below
// lds_nodes are read from the workgroup's staged copy of this tree
// (`lnodes`).
__device__ inline bool bvh_hit(const KSurface &S, const KRay &r, bool any, float &tbest,
                               int &prim, uint32_t &rec, int2 *lst,
                               const float4 *lnodes = nullptr, int lds_nodes = 0) {
  f3 inv;
  {
    const float dx = fabsf(r.dir.x) < 1e-20f ? copysignf(1e-20f, r.dir.x) : r.dir.x;
    const float dy = fabsf(r.dir.y) < 1e-20f ? copysignf(1e-20f, r.dir.y) : r.dir.y;
    const float dz = fabsf(r.dir.z) < 1e-20f ? copysignf(1e-20f, r.dir.z) : r.dir.z;
    inv = mk(1.f / dx, 1.f / dy, 1.f / dz);
  }
  int2 deep[kBvhStack - kBvhLdsStack];
  int sp = 0;
  int node = 0;
  bool found = false;
  const float inf = __builtin_inff();
  while (true) {
    float4 nd[8];
    if (node < lds_nodes) {
      const float4 *n = lnodes + 8 * node;
#pragma unroll
      for (int k = 0; k < 8; ++k) nd[k] = n[k];
    } else {
      const float4 *n = S.nodes + size_t(node) * 8u;
#pragma unroll
      for (int k = 0; k < 8; ++k) nd[k] = n[k];
    }
    const float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5];
    const float4 chf = nd[6], cnf = nd[7];
    const float lo_x[4] = {lx.x, lx.y, lx.z, lx.w}, hi_x[4] = {hx.x, hx.y, hx.z, hx.w};
    const float lo_y[4] = {ly.x, ly.y, ly.z, ly.w}, hi_y[4] = {hy.x, hy.y, hy.z, hy.w};
    const float lo_z[4] = {lz.x, lz.y, lz.z, lz.w}, hi_z[4] = {hz.x, hz.y, hz.z, hz.w};
    const int child[4] = {__float_as_int(chf.x), __float_as_int(chf.y), __float_as_int(chf.z),
                          __float_as_int(chf.w)};
    const uint32_t count[4] = {__float_as_uint(cnf.x), __float_as_uint(cnf.y),
                               __float_as_uint(cnf.z), __float_as_uint(cnf.w)};
    // entry distance of every inner child that was hit, +inf for the rest
    float ct[4];
    int cn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      ct[c] = inf;
      cn[c] = child[c];
      if (child[c] < 0) continue;
      const float tx0 = (lo_x[c] - r.ori.x) * inv.x, tx1 = (hi_x[c] - r.ori.x) * inv.x;
      const float ty0 = (lo_y[c] - r.ori.y) * inv.y, ty1 = (hi_y[c] - r.ori.y) * inv.y;
      const float tz0 = (lo_z[c] - r.ori.z) * inv.z, tz1 = (hi_z[c] - r.ori.z) * inv.z;
      const float tn = fmaxf(r.tmin, fmaxf(fminf(tx0, tx1), fmaxf(fminf(ty0, ty1), fminf(tz0, tz1))));
      const float tf = fminf(tbest, fminf(fmaxf(tx0, tx1), fminf(fmaxf(ty0, ty1), fmaxf(tz0, tz1))));
      if (!(tn <= tf * 1.0000002f + 1e-30f)) continue;
      if (count[c] == 0u) {
        ct[c] = tn;
        continue;
      }
      for (uint32_t k = 0; k < count[c]; ++k) {
        const float4 *q = S.tris + 3 * (size_t(child[c]) + k);
        const float4 a = q[0], b = q[1], d = q[2];
        float t, u, v;
        if (!tri_test(r, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(d.x, d.y, d.z), t, u, v))
          continue;
        const int p = __float_as_int(a.w);
        if (t < r.tmin || t > tbest || (found && t == tbest && p > prim)) continue;
        found = true;
        tbest = t;
        prim = p;
        rec = uint32_t(child[c]) + k;
        if (any) return true;
      }
    }
    // ascending by distance (a 5-comparator network, all in registers)
    cswap(ct[0], cn[0], ct[1], cn[1]);
    cswap(ct[2], cn[2], ct[3], cn[3]);
    cswap(ct[0], cn[0], ct[2], cn[2]);
    cswap(ct[1], cn[1], ct[3], cn[3]);
    cswap(ct[1], cn[1], ct[2], cn[2]);
    // stack the farther hits, far first, so the nearer ones are popped first
#pragma unroll
    for (int c = 3; c >= 1; --c) {
      if (!(ct[c] < inf) || sp >= kBvhStack) continue;
      const int2 e = make_int2(cn[c], __float_as_int(ct[c]));
      if (sp < kBvhLdsStack)
        lst[sp * kBvhLanes] = e;
      else
        deep[sp - kBvhLdsStack] = e;
      ++sp;
    }
    // descend into the nearest hit unless a closer triangle was found since
    // (ct[0] = +inf: no inner child was hit; the limit itself may be +inf)
    const float lim = tbest * 1.0000002f + 1e-30f;
    if (ct[0] < inf && ct[0] <= lim) {
      node = cn[0];
      continue;
    }
    // next node: skip those entered beyond the current closest hit
    node = -1;
    while (sp > 0) {
      --sp;
      const int2 e = sp < kBvhLdsStack ? lst[sp * kBvhLanes] : deep[sp - kBvhLdsStack];
      if (__int_as_float(e.y) <= tbest * 1.0000002f + 1e-30f) {
        node = e.x;
        break;
      }
    }
    if (node < 0) return found;
  }
}

// intersect(Ray, TLS) over instanced meshes (DeviceCopyableObjects.h:
// 1772-1824): the world ray walks the instance BVH (its stack, the TLS's
// bound <= kBvhStack, in private memory; the LDS stack top `lst` belongs to
// the mesh walks) and, per instance whose box it meets, meets the mesh in
// object space, ori' = A (ori + tinv), dir' = A dir, so t is a world
// distance.  A hit at t <= tbest counts when nothing was hit before
// (`found_before`: a surface hit at tbest keeps ties), else when strictly
// closer, or equally close on an instance of lower index.  Returns whether
// an instance was hit; `slot` is its position in T.inst.
__device__ inline bool tls_hit(const KTls &T, const KRay &r, bool any, bool found_before,
                               float &tbest, int &slot, int &prim, uint32_t &rec, int2 *lst) {
  f3 inv;
  {
    const float dx = fabsf(r.dir.x) < 1e-20f ? copysignf(1e-20f, r.dir.x) : r.dir.x;
    const float dy = fabsf(r.dir.y) < 1e-20f ? copysignf(1e-20f, r.dir.y) : r.dir.y;
    const float dz = fabsf(r.dir.z) < 1e-20f ? copysignf(1e-20f, r.dir.z) : r.dir.z;
    inv = mk(1.f / dx, 1.f / dy, 1.f / dz);
  }
  int2 stk[kBvhStack];
  int sp = 0;
  int node = 0;
  bool found = false;
  uint32_t best_index = 0xFFFFFFFFu;
  const float inf = __builtin_inff();
  while (true) {
    const float4 *n = T.nodes + size_t(node) * 8u;
    const float4 lx = n[0], hx = n[1], ly = n[2], hy = n[3], lz = n[4], hz = n[5];
    const float4 chf = n[6], cnf = n[7];
    const float lo_x[4] = {lx.x, lx.y, lx.z, lx.w}, hi_x[4] = {hx.x, hx.y, hx.z, hx.w};
    const float lo_y[4] = {ly.x, ly.y, ly.z, ly.w}, hi_y[4] = {hy.x, hy.y, hy.z, hy.w};
    const float lo_z[4] = {lz.x, lz.y, lz.z, lz.w}, hi_z[4] = {hz.x, hz.y, hz.z, hz.w};
    const int child[4] = {__float_as_int(chf.x), __float_as_int(chf.y), __float_as_int(chf.z),
                          __float_as_int(chf.w)};
    const uint32_t count[4] = {__float_as_uint(cnf.x), __float_as_uint(cnf.y),
                               __float_as_uint(cnf.z), __float_as_uint(cnf.w)};
    float ct[4];
    int cn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      ct[c] = inf;
      cn[c] = child[c];
      if (child[c] < 0) continue;
      const float tx0 = (lo_x[c] - r.ori.x) * inv.x, tx1 = (hi_x[c] - r.ori.x) * inv.x;
      const float ty0 = (lo_y[c] - r.ori.y) * inv.y, ty1 = (hi_y[c] - r.ori.y) * inv.y;
      const float tz0 = (lo_z[c] - r.ori.z) * inv.z, tz1 = (hi_z[c] - r.ori.z) * inv.z;
      const float tn = fmaxf(r.tmin, fmaxf(fminf(tx0, tx1), fmaxf(fminf(ty0, ty1), fminf(tz0, tz1))));
      const float tf = fminf(tbest, fminf(fmaxf(tx0, tx1), fminf(fmaxf(ty0, ty1), fmaxf(tz0, tz1))));
      if (!(tn <= tf * 1.0000002f + 1e-30f)) continue;
      if (count[c] == 0u) {
        ct[c] = tn;
        continue;
      }
      for (uint32_t k = 0; k < count[c]; ++k) {
        const KInst &I = T.inst[size_t(child[c]) + k];
        KRay lr;
        lr.ori = mulA(I.A, r.ori + mk(I.tinv[0], I.tinv[1], I.tinv[2]));
        lr.dir = mulA(I.A, r.dir);
        lr.tmin = r.tmin;
        lr.tmax = r.tmax;
        KSurface S;
        S.nodes = I.nodes;
        S.tris = I.tris;
        float t = tbest;
        int p = -1;
        uint32_t rc = 0;
        if (!bvh_hit(S, lr, any, t, p, rc, lst)) continue;
        if ((found_before || found) && !(t < tbest || (found && t == tbest && I.index < best_index)))
          continue;
        found = true;
        tbest = t;
        slot = child[c] + int(k);
        prim = p;
        rec = rc;
        best_index = I.index;
        if (any) return true;
      }
    }
    cswap(ct[0], cn[0], ct[1], cn[1]);
    cswap(ct[2], cn[2], ct[3], cn[3]);
    cswap(ct[0], cn[0], ct[2], cn[2]);
    cswap(ct[1], cn[1], ct[3], cn[3]);
    cswap(ct[1], cn[1], ct[2], cn[2]);
#pragma unroll
    for (int c = 3; c >= 1; --c) {
      if (!(ct[c] < inf) || sp >= kBvhStack) continue;
      stk[sp++] = make_int2(cn[c], __float_as_int(ct[c]));
    }
    const float lim = tbest * 1.0000002f + 1e-30f;
    if (ct[0] < inf && ct[0] <= lim) {
      node = cn[0];
      continue;
    }
    node = -1;
    while (sp > 0) {
      const int2 e = stk[--sp];
      if (__int_as_float(e.y) <= tbest * 1.0000002f + 1e-30f) {
        node = e.x;
        break;
      }
    }
    if (node < 0) return found;
  }
}

// The matte triangle hit by a pass: surface `id` (< num_surfaces) or
// instance id - num_surfaces of the TLS, with the instance's normal
// transform (identity: nxfm null)
struct SurfRef {
  const float4 *tris;
  f3 color;
  uint32_t obj_id, inst_id;
  const float *nxfm;
};

__device__ inline SurfRef surface_ref(const KScene &sc, int id) {
  SurfRef R;
  if (id < sc.num_surfaces) {
    KSurface S = sc.surf[0];
    for (int j = 1; j < sc.num_surfaces; ++j) {
      const int i = uidx(j);
      const KSurface C = sc.surf[i];
      const bool m = i == id;
      S.tris = m ? C.tris : S.tris;
      S.color = m ? C.color : S.color;
      S.obj_id = m ? C.obj_id : S.obj_id;
      S.inst_id = m ? C.inst_id : S.inst_id;
    }
    R.tris = S.tris;
    R.color = S.color;
    R.obj_id = S.obj_id;
    R.inst_id = S.inst_id;
    R.nxfm = nullptr;
  } else {
    const KInst &I = sc.tls.inst[id - sc.num_surfaces];
    R.tris = I.tris;
    R.color = mk(I.color[0], I.color[1], I.color[2]);
    R.obj_id = I.obj_id;
    R.inst_id = I.inst_id;
    R.nxfm = I.N;
  }
  return R;
}

// the geometric normal of triangle record `rec` in world space:
// normalize(nxfm * normalize(e1 x e2)) (getNormals, common.h:233-264, then
// Pathtrace_impl.cpp:243-247; identity: normalised twice)
__device__ inline f3 surface_normal(const SurfRef &R, uint32_t rec) {
  const float4 e1 = R.tris[3 * size_t(rec) + 1], e2 = R.tris[3 * size_t(rec) + 2];
  const f3 n = normalize(cross(mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z)));
  return normalize(R.nxfm ? mulA(R.nxfm, n) : n);
}

// intersect(Ray, Quad) (DeviceCopyableObjects.h:2209-2230): the closer of
// the two triangles, t > 0, no [tmin, tmax] test (intersectLights)
__device__ inline bool quad_hit(const KLight &L, const KRay &r, float &t) {
  const f3 d = L.e1 + L.e2;
  float t1, t2, u, v;
  const bool h1 = tri_test(r, L.v1, L.e1, d, t1, u, v) && t1 > 0.f;
  const bool h2 = tri_test(r, L.v1, d, L.e2, t2, u, v) && t2 > 0.f;
  if (!h1 && !h2) return false;
  t = h1 ? (h2 && t2 < t1 ? t2 : t1) : t2;
  return true;
}

// power_heuristic (visionaray, parity unpinned): f^2 / (f^2 + g^2), 0 when
// both vanish
__device__ inline float power_heuristic(float f, float g) {
  const float d = f * f + g * g;
  return d > 0.f ? (f * f) / d : 0.f;
}

// ---------------------------------------------------------------------------
// Woodcock / delta tracking through one volume (DeviceCopyableObjects.h:
// 680-796) with the macrocell DDA of renderer/DDA.h:14-112.
// ---------------------------------------------------------------------------
struct VolHit {
  float t;
  f3 albedo;
  float extinction;  // classified opacity at the collision (HitRecordVolume::extinction)
  uint32_t rec;      // sample_field's record of the collision
  bool hit;
};

// Resumable delta tracking through one volume.  The walk is split into
// trav_begin (box clip, voxel space, DDA set-up) and trav_step (one step of a
// flat loop over "set up the current cell" and "one free-flight step in
// it"), so that a persistent wave can hand lanes whose ray ended new rays
// while the other lanes keep walking.  Per cell the draws are those of the
// reference's callback: none for a zero majorant, one for an empty [t0, t1)
// range (dda3 keeps calling it with empty ranges past tmax), else one per
// free-flight step plus one per tentative collision.
struct Trav {
  f3 ori, dir;      // voxel space, unit direction
  float tmin;       // voxel-space start; the DDA runs on t - tmin
  float tm;         // grid: tmax - tmin; no grid: tmax
  float invUD, dt_scale;
  int cell[3];
  float dist[3], tnext[3];
  float t0, t1;     // current cell's range relative to tmin
  float t, tend;    // free-flight position and the cell's end (absolute)
  float maj, scale;
  int cells_left;   // < 0: no majorant grid, one range with majorant 1
  bool in_cell;
  float tr;         // ratio tracking: the transmittance estimate so far
};

// false: the ray misses the volume's box
__device__ inline bool trav_begin(const KVolume &v, KRay ray, Trav &T) {
  BoxHit bh = intersect_box(ray, v.blo, v.bhi);
  if (!bh.hit) return false;
  ray.tmin = fmaxf(ray.tmin, bh.tnear);
  ray.tmax = fminf(ray.tmax, bh.tfar);
  const f3 ori = v.S * (ray.ori + v.T);
  f3 dir = v.S * ray.dir;
  T.dt_scale = length(dir);
  dir = normalize(dir);
  T.ori = ori;
  T.dir = dir;
  T.tmin = ray.tmin * T.dt_scale;
  const float tmax = ray.tmax * T.dt_scale;
  T.invUD = 1.f / (v.unit_distance * T.dt_scale);
  T.t0 = 0.f;
  T.in_cell = false;
  T.tr = 1.f;
  if (v.gdims[0] == 0) {
    T.tm = tmax;
    T.cells_left = -1;
    return true;
  }
  // dda3 with the ray moved so that tmin becomes 0 (DDA.h:21-25)
  const f3 o = ori + dir * T.tmin;
  T.tm = tmax - T.tmin;
  const f3 rcp = mk(rcp_dir(dir.x), rcp_dir(dir.y), rcp_dir(dir.z));
  // grid bounds in voxel space: [0,1]^3 for structured fields, the world
  // bounds for nanovdb (identity voxel transform)
  const f3 l = (v.glo - o) * rcp;
  const f3 h = (v.ghi - o) * rcp;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int G = v.gdims[a];
    const float oa = comp(o, a), da = comp(dir, a);
    const float ga = comp(v.glo, a);
    float c = (oa - ga) / (comp(v.ghi, a) - ga) * float(G);  // projectOnGrid
    c = fminf(fmaxf(c, -1.f), float(G));
    T.cell[a] = min(max(int(c), 0), G - 1);
    const float la = comp(l, a), ha = comp(h, a);
    const float tn = fminf(la, ha), tf = fmaxf(la, ha);
    if (da == 0.f) {
      T.dist[a] = 0.f;
      T.tnext[a] = __builtin_inff();
    } else {
      T.dist[a] = fmaxf(0.f, (tf - tn) / float(G));
      T.tnext[a] = da > 0.f ? tn + float(T.cell[a] + 1) * T.dist[a]
                            : tn + float(G - T.cell[a]) * T.dist[a];
    }
  }
  // a ray visits at most gx+gy+gz cells; the cap only guards degenerate
  // (non-finite) rays
  T.cells_left = v.gdims[0] + v.gdims[1] + v.gdims[2] + 3;
  return true;
}

// One step of the walk.  Returns true when it ended: on a collision (`hit`,
// voxel-space `t_hit`, classified sample `smp`) or on leaving the grid.
// `mb`/`mstep`: the volume's majorant bytes in LDS and their step (mb null:
// the float majorants in global memory).  RATIO: ratio tracking -- a
// tentative collision multiplies T.tr by 1 - alpha/majorant and the walk goes
// on (no acceptance draw), so it ends only on leaving the grid.
template <bool ANYF = true, bool RATIO = false>
__device__ inline bool trav_step(const KVolume &v, const float4 *lds, const uint8_t *mb,
                                 float mstep, Trav &T, uint32_t &rs, uint32_t &lookups,
                                 bool &hit, float &t_hit, float4 &smp, uint32_t &rec_hit) {
  bool advance = false;
  if (!T.in_cell) {
    float a0, a1, maj;
    if (T.cells_left < 0) {
      a0 = T.tmin;
      a1 = T.tm;
      maj = 1.f;
    } else {
      const float tc = fminf(T.tnext[0], fminf(T.tnext[1], T.tnext[2]));
      T.t1 = fminf(tc, T.tm);
      a0 = T.tmin + T.t0;
      a1 = T.tmin + T.t1;
      const int li = (T.cell[2] * v.gdims[1] + T.cell[1]) * v.gdims[0] + T.cell[0];
      maj = mb ? float(mb[li]) * mstep : v.majorants[li];
    }
    if (maj > 0.f) {
      if (a0 < a1) {
        T.maj = maj;
        T.scale = 1.f / (maj * T.invUD);
        T.t = a0;
        T.tend = a1;
        T.in_cell = true;
      } else {
        rs = 1664525u * rs + 1013904223u;  // the step from an empty range
        advance = true;
      }
    } else {
      advance = true;
    }
  }
  if (T.in_cell) {
    T.t -= __logf(1.f - rnd(rs)) * T.scale;
    if (!(T.t < T.tend)) {  // also leaves on NaN
      T.in_cell = false;
      advance = true;
    } else {
      float val;
      uint32_t rec;
      ++lookups;
      // no acceptance draw where the field has no value
      // (DeviceCopyableObjects.h:739-753)
      if (sample_field<ANYF>(v, T.ori + T.dir * T.t, val, rec)) {
        if (RATIO) {
          T.tr = T.tr * (1.f - classify_alpha(v, lds, val) / T.maj);
        } else if (classify_alpha(v, lds, val) >= rnd(rs) * T.maj) {
          // colour only for an accepted collision
          hit = true;
          t_hit = T.t;
          smp = classify(v, lds, val);
          rec_hit = rec;
          return true;
        }
      }
    }
  }
  if (advance) {
    if (T.cells_left < 0) return true;
    const float tc = fminf(T.tnext[0], fminf(T.tnext[1], T.tnext[2]));
    bool out_of_grid = false;
    if (T.tnext[0] == tc) {
      T.tnext[0] += T.dist[0];
      T.cell[0] += T.dir.x > 0.f ? 1 : -1;
      out_of_grid = T.cell[0] == (T.dir.x > 0.f ? v.gdims[0] : -1);
    }
    if (!out_of_grid && T.tnext[1] == tc) {
      T.tnext[1] += T.dist[1];
      T.cell[1] += T.dir.y > 0.f ? 1 : -1;
      out_of_grid = T.cell[1] == (T.dir.y > 0.f ? v.gdims[1] : -1);
    }
    if (!out_of_grid && T.tnext[2] == tc) {
      T.tnext[2] += T.dist[2];
      T.cell[2] += T.dir.z > 0.f ? 1 : -1;
      out_of_grid = T.cell[2] == (T.dir.z > 0.f ? v.gdims[2] : -1);
    }
    if (out_of_grid || --T.cells_left == 0) return true;
    T.t0 = T.t1;
  }
  return false;
}

__device__ inline VolHit make_hit(bool hit, float t_hit, float4 smp, float dt_scale,
                                  uint32_t rec) {
  VolHit out;
  out.hit = hit;
  out.rec = rec;
  out.t = hit ? t_hit / dt_scale : FLT_MAX;
  out.albedo = mk(smp.x, smp.y, smp.z);
  out.extinction = smp.w;
  return out;
}

// Ratio tracking of a shadow / AO ray through one volume (an extension: the
// same expectation as the binary transmittance of delta tracking): the
// estimate Tr in [0, 1]
template <bool ANYF = true>
__device__ inline float ratio_track(const KVolume &v, const float4 *lds, const uint8_t *mb,
                                    float mstep, KRay ray, uint32_t &rs, uint32_t &lookups) {
  Trav T;
  bool hit = false;
  float t_hit = 0.f;
  float4 smp;
  uint32_t rec;
  if (!trav_begin(v, ray, T)) return 1.f;
  while (!trav_step<ANYF, true>(v, lds, mb, mstep, T, rs, lookups, hit, t_hit, smp, rec)) {
  }
  return T.tr;
}

// intersect(Ray, Volume) (DeviceCopyableObjects.h:680-796) in one call
template <bool ANYF = true>
__device__ inline VolHit delta_track(const KVolume &v, const float4 *lds, const uint8_t *mb,
                                     float mstep, KRay ray, uint32_t &rs, uint32_t &lookups) {
  Trav T;
  bool hit = false;
  float t_hit = 0.f;
  float4 smp = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t rec = kNoRec;
  if (!trav_begin(v, ray, T)) return make_hit(false, 0.f, smp, 1.f, rec);
  while (!trav_step<ANYF>(v, lds, mb, mstep, T, rs, lookups, hit, t_hit, smp, rec)) {
  }
  return make_hit(hit, t_hit, smp, T.dt_scale, rec);
}

// ---------------------------------------------------------------------------
// Sampling helpers
// ---------------------------------------------------------------------------
// sin/cos of an angle given in revolutions (turns).  The hardware sine and
// cosine take revolutions, so the sampled angles are formed in turns.
__device__ inline void sincos_turns(float t, float &s, float &c) {
#if VRT_NATIVE_TRIG
  s = __builtin_amdgcn_sinf(t);
  c = __builtin_amdgcn_cosf(t);
#else
  sincosf(6.28318530717958647692f * t, &s, &c);
#endif
}

__device__ inline f3 uniform_sample_sphere(float u1, float u2) {
  float z = 1.f - 2.f * u1;
  float r = sqrtf(fmaxf(0.f, 1.f - z * z));
  float s, c;
  sincos_turns(u2, s, c);  // phi = 2 pi u2
  return mk(r * c, r * s, z);
}
// Shirley and Chiu's concentric map of [0,1)^2 onto the unit disk
__device__ inline void concentric_disk(float u1, float u2, float &dx, float &dy) {
  float sx = 2.f * u1 - 1.f, sy = 2.f * u2 - 1.f;
  dx = 0.f, dy = 0.f;
  if (!(sx == 0.f && sy == 0.f)) {
    // theta = pi/4 (sy/sx) or pi/2 - pi/4 (sx/sy), in turns
    float r, turns;
    if (fabsf(sx) > fabsf(sy)) {
      r = sx;
      turns = 0.125f * (sy / sx);
    } else {
      r = sy;
      turns = 0.25f - 0.125f * (sx / sy);
    }
    float s, c;
    sincos_turns(turns, s, c);
    dx = r * c;
    dy = r * s;
  }
}
__device__ inline f3 cosine_sample_hemisphere(float u1, float u2) {
  float dx, dy;
  concentric_disk(u1, u2, dx, dy);
  return mk(dx, dy, sqrtf(fmaxf(0.f, 1.f - dx * dx - dy * dy)));
}

// Thin lens (the perspective camera's apertureRadius / focusDistance,
// camera/Perspective.cpp:42-43): the pinhole ray's point on the plane in
// focus, focus_distance along the view direction, seen from the lens point
// off = lens radius * concentric_disk(u1, u2) in the plane of U and V
// (Pharr, Jakob and Humphreys, PBRT 3rd ed., 6.2.3); the new direction
// dir * ft - off is that point minus the new origin ori + off
__device__ inline void thin_lens(f3 &ori, f3 &dir, const KCamera &c, float u1, float u2) {
  float dx, dy;
  concentric_disk(u1, u2, dx, dy);
  const float ft = c.focus_distance / dot(dir, c.W);
  const f3 off = c.U * (c.lens_u * dx) + c.V * (c.lens_v * dy);
  dir = normalize(dir * ft - off);
  ori = ori + off;
}
__device__ inline f3 faceforward(f3 n, f3 i, f3 nref) { return dot(i, nref) < 0.f ? -n : n; }

// Henyey-Greenstein phase function (an extension: the reference's phase is
// isotropic).  hg_4pi = 4 pi p_HG(cos theta) = (1 - g^2) / (1 + g^2 - 2 g
// cos theta)^(3/2), exactly 1 at g = 0; cos theta is taken between the ray
// direction and the scattered (or light) direction, so g > 0 scatters
// forward.
__device__ inline float hg_4pi(float g, float cosT) {
  const float d = 1.f + g * g - 2.f * g * cosT;
  return (1.f - g * g) / (d * sqrtf(d));
}
// Direction with density p_HG around unit d: cos theta from the inverse CDF
// (1 + g^2 - ((1 - g^2) / (1 - g + 2 g u1))^2) / (2 g), phi = 2 pi u2, in
// the orthonormal basis of Duff et al. (2017) around d.  g != 0.
__device__ inline f3 hg_sample(float g, f3 d, float u1, float u2) {
  const float s = (1.f - g * g) / (1.f - g + 2.f * g * u1);
  const float cosT = fminf(fmaxf((1.f + g * g - s * s) / (2.f * g), -1.f), 1.f);
  const float sinT = sqrtf(fmaxf(0.f, 1.f - cosT * cosT));
  float sp, cp;
  sincos_turns(u2, sp, cp);
  const float sign = copysignf(1.f, d.z);
  const float a = -1.f / (sign + d.z);
  const float b = d.x * d.y * a;
  const f3 t1 = mk(1.f + sign * d.x * d.x * a, sign * b, -sign * d.x);
  const f3 t2 = mk(b, sign + d.y * d.y * a, -d.y);
  return t1 * (sinT * cp) + t2 * (sinT * sp) + d * cosT;
}

// ---------------------------------------------------------------------------
// Frame write: Frame::writeSample (DeviceCopyableObjects.h:2776-2781)
// ---------------------------------------------------------------------------
__device__ inline uint32_t cvt8(float f) { return uint32_t(255.f * fminf(fmaxf(f, 0.f), 1.f)); }
__device__ inline float linear_to_srgb(float x) {
  return x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
}

__device__ inline void tone_map(const KFrame &F, size_t idx, float4 c);

__device__ inline void write_sample(const KFrame &F, int x, int y, int accumID, float4 c,
                                    float depth, bool stochastic) {
  const size_t idx = size_t(y) * F.width + x;
  if (F.depth && depth <= F.depth[idx]) F.depth[idx] = depth;
  if (F.accum64) {
    double *a = F.accum64 + 4 * idx;
    a[0] += double(c.x);
    a[1] += double(c.y);
    a[2] += double(c.z);
    a[3] += double(c.w);
  }
  if (stochastic && F.accum) {
    // running mean, alpha = 1/(accumID+1) (DeviceCopyableObjects.h:2713-2716)
    const float alpha = 1.f / float(accumID + 1);
    const float4 prev = F.accum[idx];
    c = make_float4((1.f - alpha) * prev.x + alpha * c.x, (1.f - alpha) * prev.y + alpha * c.y,
                    (1.f - alpha) * prev.z + alpha * c.z, (1.f - alpha) * prev.w + alpha * c.w);
    F.accum[idx] = c;
  }
  tone_map(F, idx, c);
}

// toneMap (DeviceCopyableObjects.h:2720-2739), cvt_uint32 (frame/common.h:31-47)
__device__ inline void tone_map(const KFrame &F, size_t idx, float4 c) {
  if (F.color_type == VRT_TYPE_FLOAT32_VEC4) {
    reinterpret_cast<float4 *>(F.color)[idx] = c;
  } else if (F.color_type == VRT_TYPE_UFIXED8_VEC4) {
    reinterpret_cast<uint32_t *>(F.color)[idx] =
        cvt8(c.x) | (cvt8(c.y) << 8) | (cvt8(c.z) << 16) | (cvt8(c.w) << 24);
  } else if (F.color_type == VRT_TYPE_UFIXED8_RGBA_SRGB) {
    reinterpret_cast<uint32_t *>(F.color)[idx] = cvt8(linear_to_srgb(c.x)) |
                                                 (cvt8(linear_to_srgb(c.y)) << 8) |
                                                 (cvt8(linear_to_srgb(c.z)) << 16) |
                                                 (cvt8(c.w) << 24);
  }
}

__device__ inline void wave_add_counter(unsigned long long *ctr, uint32_t v) {
  // 64-lane butterfly, then one atomic per wave
  unsigned long long s = v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(ctr, s);
}

// ---------------------------------------------------------------------------
// Raycast: Raycast_impl.cpp:245-322 + rayMarchVolume<0>
// (VolumeIntegration.h:374-470)
// ---------------------------------------------------------------------------
// The march fetches kRaySteps steps ahead: their positions are known before
// any of them is composited, so their voxel loads are in flight together;
// compositing then runs step by step with the reference's tests (t <= tmax,
// alpha < 0.99), and fetches past the end are discarded.
constexpr int kRaySteps = 4;

// ---------------------------------------------------------------------------
// HDRI environment light (sampleCDF.h, DeviceCopyableObjects.h:2398-2418,
// HDRI.cpp).  The mappings between directions and map coordinates keep the
// reference's arithmetic: M_PI terms in double, rounded to float once.
// ---------------------------------------------------------------------------
constexpr double kPiD = 3.14159265358979323846;

// upper_bound: the first entry greater than x (sampleCDF.h:14-31)
__device__ inline uint32_t cdf_upper_bound(const float *d, uint32_t n, float x) {
  uint32_t first = 0, count = n;
  while (count > 0) {
    const uint32_t step = count / 2, it = first + step;
    if (!(x < d[it])) {
      first = it + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  return first;
}

// sample_cdf (sampleCDF.h:40-54): the bin of x and its probability; a bin
// past the end (a map or row summing to 0) has probability 0 here, the
// reference leaves it unset
__device__ inline uint32_t cdf_bin(const float *d, uint32_t n, float x, float &pdf) {
  const uint32_t i = cdf_upper_bound(d, n, x);
  pdf = i == 0 ? d[0] : (i < n ? d[i] - d[i - 1] : 0.f);
  return i;
}

// sampleCDF (sampleCDF.h:88-98): the row from the row totals, then the
// column from that row
__device__ inline void hdri_cdf(const KHdri &H, float rx, float ry, uint32_t &x, uint32_t &y,
                                float &pdfx, float &pdfy) {
  y = min(cdf_bin(H.rows, uint32_t(H.h), ry, pdfy), uint32_t(H.h - 1));
  x = cdf_bin(H.cols + size_t(y) * uint32_t(H.w), uint32_t(H.w), rx, pdfx);
}

// toPolar (sampleCDF.h:58-72)
__device__ inline f3 hdri_to_polar(float u, float v) {
  const float theta = float(2.0 * kPiD * double(u) + -kPiD / 2.0);
  const float phi = float(kPiD * double(v));
  const float sp = sinf(phi);
  return mk(-(cosf(theta) * sp), cosf(phi), sinf(theta) * sp);
}

// toUV (sampleCDF.h:74-85)
__device__ inline void hdri_to_uv(f3 n, float &u, float &v) {
  const float a = atan2f(n.x, n.z);
  u = float((double(a) + kPiD / 2.0) / (kPiD * 2.0) + kPiD * double(28.670f / 360.0f));
  v = fminf(fmaxf(float(double(acosf(n.y)) / kPiD), .001f), .999f);
}

// intensity (DeviceCopyableObjects.h:2413-2417): the map at toUV of the
// direction in the light's frame, bilinear with clamp addressing at texel
// centres, times scale
__device__ inline f3 hdri_radiance(const KHdri &H, f3 dir) {
  float u, v;
  hdri_to_uv(mulA(H.to_local, dir), u, v);
  int i[2], j[2];
  float wx, wy;
  {
    float c = fminf(fmaxf(u * float(H.w) - 0.5f, -1.f), float(H.w));
    const float fl = floorf(c);
    wx = c - fl;
    i[0] = min(max(int(fl), 0), H.w - 1);
    i[1] = min(max(int(fl) + 1, 0), H.w - 1);
    c = fminf(fmaxf(v * float(H.h) - 0.5f, -1.f), float(H.h));
    const float fy = floorf(c);
    wy = c - fy;
    j[0] = min(max(int(fy), 0), H.h - 1);
    j[1] = min(max(int(fy) + 1, 0), H.h - 1);
  }
  const float4 t00 = H.tex[size_t(j[0]) * H.w + i[0]], t10 = H.tex[size_t(j[0]) * H.w + i[1]];
  const float4 t01 = H.tex[size_t(j[1]) * H.w + i[0]], t11 = H.tex[size_t(j[1]) * H.w + i[1]];
  auto lp = [](float a, float b, float t) { return a + t * (b - a); };
  const f3 r = mk(lp(lp(t00.x, t10.x, wx), lp(t01.x, t11.x, wx), wy),
                  lp(lp(t00.y, t10.y, wx), lp(t01.y, t11.y, wx), wy),
                  lp(lp(t00.z, t10.z, wx), lp(t01.z, t11.z, wx), wy));
  return r * H.scale;
}

// the light pdf of a direction that hit the map (Pathtrace_impl.cpp:160-172):
// the map coordinates run through the CDF as if they were the random numbers
__device__ inline float hdri_pdf(const KHdri &H, f3 raydir) {
  const f3 d = mulA(H.to_local, raydir);
  float u, v;
  hdri_to_uv(d, u, v);
  uint32_t x, y;
  float pdfx, pdfy;
  hdri_cdf(H, u, v, x, y, pdfx, pdfy);
  const float sinT = sinf(acosf(fminf(fmaxf(d.y, -1.f), 1.f)));
  if (sinT == 0.f) return 0.f;
  const float pi = 3.14159265358979323846f;
  return (pdfx * pdfy) * float(uint32_t(H.w) * uint32_t(H.h)) / (2.0f * pi * pi * sinT);
}

// sampleLight (renderer/common.h:1432-1466).  Directional: the stored
// direction, pdf 1, distance FLT_MAX, no draws.  Quad (area_light<Quad>, not
// vendored: parity unpinned): one draw picks a triangle
// (DeviceCopyableObjects.h:2258-2270), two place a uniform point on it;
// pdf = 1 / (A |cos theta_light|), so that with the 1/dist^2 of sampleLight
// the NEE estimate is the area-sampling estimator.
struct LightSmp {
  f3 intensity, dir;
  float pdf, dist, dist2;
  int type;
};

template <bool AREA>
__device__ inline LightSmp sample_light(const KScene &sc, const KLight &L, f3 hitPos,
                                       uint32_t &rs) {
  LightSmp ls;
  ls.type = L.type;
  ls.intensity = L.intensity;
  if (AREA && L.type == VRT_LIGHT_HDRI) {
    // asHDRI.sample: a cell by the CDFs (rx then ry, drawn left to right),
    // the direction of its corner, pdf = pdfx pdfy w h / 4 pi; infinitely far
    const KHdri &H = sc.hdri;
    const float rx = rnd(rs), ry = rnd(rs);
    uint32_t x, y;
    float pdfx, pdfy;
    hdri_cdf(H, rx, ry, x, y, pdfx, pdfy);
    const float invjac = float(uint32_t(H.w) * uint32_t(H.h)) / float(4.0 * kPiD);
    ls.dir = mulA(H.to_world, hdri_to_polar(float(x) / float(H.w), float(y) / float(H.h)));
    ls.intensity = hdri_radiance(H, ls.dir);
    ls.dist = FLT_MAX;
    ls.pdf = pdfx * pdfy * invjac;
    ls.dist2 = 1.f;
  } else if (AREA && L.type == VRT_LIGHT_QUAD) {
    const float r = rnd(rs);
    const f3 d = L.e1 + L.e2;
    const bool first = L.a1 / (L.a1 + L.a2) < r;
    const f3 te1 = first ? L.e1 : d, te2 = first ? d : L.e2;
    const float u1 = rnd(rs), u2 = rnd(rs);
    const float su = sqrtf(u1);
    const f3 v2 = L.v1 + te1, v3 = L.v1 + te2;
    const f3 pos = L.v1 * (1.f - su) + v2 * (su * (1.f - u2)) + v3 * (su * u2);
    ls.dir = pos - hitPos;
    ls.dist = length(ls.dir);
    const float cosl = fabsf(dot(L.nrm, ls.dir * (1.f / ls.dist)));
    ls.pdf = cosl > 0.f ? 1.f / ((L.a1 + L.a2) * cosl) : 0.f;
    ls.dist2 = ls.dist * ls.dist;
  } else {
    ls.dir = L.dir;
    ls.dist = FLT_MAX;
    ls.pdf = 1.f;
    ls.dist2 = 1.f;
  }
  return ls;
}

__device__ inline float safe_rcp(float f) { return f > 0.f ? 1.f / f : 0.f; }

// The ray caster's backdrop (Raycast_impl.cpp:158-173): the closest visible
// light the primary ray meets -- a quad, or the environment at FLT_MAX -- as
// float4(intensity, 1), else the background colour
__device__ inline float4 light_backdrop(const KScene &sc, const KRay &ray, float4 bg) {
  bool lh = false;
  float lt = FLT_MAX;
  int lid = -1;
  for (int i = 0; i < sc.num_lights; ++i) {
    const KLight &L = sc.light[i];
    if (!L.visible) continue;
    if (L.type == VRT_LIGHT_HDRI) {
      if (!lh) {
        lh = true;
        lid = i;
      }
      continue;
    }
    if (L.type != VRT_LIGHT_QUAD) continue;
    float t;
    if (quad_hit(L, ray, t) && t < lt) {
      lh = true;
      lt = t;
      lid = i;
    }
  }
  if (!lh) return bg;
  f3 I = mk(0.f);
  for (int j = 0; j < sc.num_lights; ++j) {
    const int i = uidx(j);
    if (i == lid) I = sc.light[i].type == VRT_LIGHT_HDRI ? hdri_radiance(sc.hdri, ray.dir)
                                                        : sc.light[i].intensity;
  }
  return make_float4(I.x, I.y, I.z, 1.f);
}

// Gradient shading of one ray-caster sample (VolumeIntegration.h:420-459):
// matte (albedo / pi, common.h:825-828, 1033-1048) under every light, along
// the faceforwarded gradient normal at the voxel-space sample point P, plus
// the ambient term.  Lights are sampled at P itself, in voxel space, as the
// reference does.
template <bool ANYF>
__device__ inline f3 gradient_shade(const KScene &sc, const KVolume &v, const KParams &P, f3 Pv,
                                    f3 viewDir, f3 albedo, uint32_t &rs) {
  f3 gn = gradient_normal<ANYF>(v, Pv);
  gn = faceforward(gn, viewDir, gn);
  f3 shaded = mk(0.f);
  if (length(gn) > 1e-10f) {
    const float inv_pi = 0.318309886183790671538f;
    for (int i = 0; i < sc.num_lights; ++i) {
      const LightSmp ls = sample_light<true>(sc, sc.light[i], Pv, rs);
      const f3 lightDir = normalize(ls.dir);
      const f3 lightIntensity = ls.intensity * safe_rcp(ls.dist2);
      const float NdotL = fmaxf(0.f, dot(gn, lightDir));
      shaded = shaded + albedo * inv_pi * lightIntensity * NdotL * safe_rcp(ls.pdf);
    }
  }
  return shaded + albedo * P.ambient_color * P.ambient_radiance;
}

// Frame::fillGBuffer (DeviceCopyableObjects.h:2745-2774): the channels other
// than depth take the values of the pixel's closest sample (first of equal
// depths, Pathtrace_impl.cpp:616) when the frame has a depth channel and that
// depth is <= the stored one.  The path tracer writes them at a sample's
// first hit when it is closer than the pixel's earlier samples and passes
// the depth test, which leaves the same values: the last write is the
// closest sample, and it passes exactly when the reference's test does.
__device__ inline bool gbuf_wanted(const KFrame &F) {
  return F.depth && (F.normal || F.albedo || F.prim_id || F.obj_id || F.inst_id);
}

__device__ inline void gbuf_write(const KFrame &F, size_t i, f3 n, f3 a, uint32_t prim,
                                  uint32_t obj, uint32_t inst) {
  if (F.normal) {
    F.normal[3 * i] = n.x;
    F.normal[3 * i + 1] = n.y;
    F.normal[3 * i + 2] = n.z;
  }
  if (F.albedo) {
    F.albedo[3 * i] = a.x;
    F.albedo[3 * i + 1] = a.y;
    F.albedo[3 * i + 2] = a.z;
  }
  if (F.prim_id) F.prim_id[i] = prim;
  if (F.obj_id) F.obj_id[i] = obj;
  if (F.inst_id) F.inst_id[i] = inst;
}

// a pixel none of whose samples hit a volume or surface: the default
// PixelSample (frame/common.h:15-26: ids ~0u; normal and albedo, left
// unset there, written as 0) at depth 1e31
__device__ inline void gbuf_miss(const KFrame &F, size_t i) {
  if (1e31f <= F.depth[i]) gbuf_write(F, i, mk(0.f), mk(0.f), ~0u, ~0u, ~0u);
}

// One pixel of the ray caster (Raycast_impl.cpp).  SURF: the scene has
// surfaces (the BVH walk and surface shading are compiled in, with the
// traversal stack's top in LDS after the tables)
template <bool ANYF, bool SHADE, bool SURF, bool INST>
__device__ inline void raycast_pixel(const KScene &sc, const KCamera &cam, const KParams &P,
                                     const KFrame &F, int x, int y, float4 *s_tf,
                                     const float2 *s_pow, uint32_t &lookups) {
  const bool active = x < F.width && y < F.height;
  if (active) {
    float4 result = P.bg;
    int hit_vol = -1;
    // g-buffer of the sample (Raycast_impl.cpp:18-20, 135-143, 230-237)
    float gb_depth = 1e31f;
    f3 gb_n = mk(0.f), gb_alb = mk(0.f);
    uint32_t gb_prim = ~0u, gb_obj = ~0u, gb_inst = ~0u;
    // light samples (surface and gradient shading) from the pixel's
    // path-tracer stream: the reference's ray caster leaves its generator
    // unseeded; a thin lens's sample, if any, comes first
    uint32_t rs = (SHADE || SURF || cam.lens_u > 0.f) ? tea4(uint32_t(y * F.width + x), 0u)
                                                           : 0u;
    if (sc.num_volumes > 0 || sc.num_surfaces > 0 || sc.tls.num_inst > 0) {  // an empty scene shows the background
      KRay ray = primary_ray(cam, float(x), float(y), float(F.width), float(F.height));
      if (cam.lens_u > 0.f) {
        const float u1 = rnd(rs);
        const float u2 = rnd(rs);
        thin_lens(ray.ori, ray.dir, cam, u1, u2);
      }
      clip_ray(ray, P);
      float4 surf = make_float4(0.f, 0.f, 0.f, 0.f);  // surfaceColor, surfaceAlpha
      if constexpr (SURF) {
        // the closest surface (Raycast_impl.cpp:27-156): matte under every
        // light along the surface normal (not faceforwarded) plus ambient;
        // matte is opaque, so the first hit ends the surface loop and the
        // volume march stops there
        int2 *bstack = reinterpret_cast<int2 *>(reinterpret_cast<char *>(s_tf) +
                                                scene_lds_bytes(sc)) + threadIdx.x;
        float st = ray.tmax;
        int sid = -1, prim = -1;
        uint32_t srec = 0;
        for (int i = 0; i < sc.num_surfaces; ++i) {
          float t = st;
          uint32_t rec;
          int p;
          if (bvh_hit(sc.surf[i], ray, false, t, p, rec, bstack,
                      staged_nodes(sc, s_tf) + sc.staged[i].lds_first, sc.staged[i].lds_nodes)) {
            st = t;
            sid = i;
            srec = rec;
            prim = p;
          }
        }
        if (INST && sc.tls.num_inst > 0) {
          float t = st;
          int slot = -1, p = -1;
          uint32_t rec = 0;
          if (tls_hit(sc.tls, ray, false, sid >= 0, t, slot, p, rec, bstack)) {
            st = t;
            sid = sc.num_surfaces + slot;
            srec = rec;
            prim = p;
          }
        }
        if (sid >= 0) {
          const SurfRef S = surface_ref(sc, sid);
          const float4 v1 = S.tris[3 * size_t(srec)];
          const f3 hitPos = ray.ori + ray.dir * st;
          // getNormals, then normalize(nxfm * n) under the instance
          const f3 sn = surface_normal(S, srec);
          const float inv_pi = 0.318309886183790671538f;
          f3 shaded = mk(0.f);
          for (int i = 0; i < sc.num_lights; ++i) {
            const LightSmp ls = sample_light<true>(sc, sc.light[i], hitPos, rs);
            const f3 lightDir = normalize(ls.dir);
            const f3 lightIntensity = ls.intensity * safe_rcp(ls.dist2);
            const float NdotL = fmaxf(0.f, dot(sn, lightDir));
            shaded = shaded + S.color * inv_pi * lightIntensity * NdotL * safe_rcp(ls.pdf);
          }
          shaded = shaded + S.color * P.ambient_color * P.ambient_radiance;
          surf = make_float4(shaded.x, shaded.y, shaded.z, 1.f);
          ray.tmax = st;
          gb_depth = st;
          gb_n = sn;
          gb_alb = S.color;
          gb_prim = __float_as_uint(v1.w);
          gb_obj = S.obj_id;
          gb_inst = S.inst_id;
          (void)prim;
        }
      }
      // over(float4(surfaceColor, surfaceAlpha), backdrop) (:167-173)
      {
        const float4 b = light_backdrop(sc, ray, P.bg);
        const float a = 1.f - surf.w;
        result = make_float4(surf.x + a * b.x, surf.y + a * b.y, surf.z + a * b.z, surf.w + a * b.w);
      }
      // closest volume box (intersectVolumeBounds, DeviceCopyableObjects.h:1841-1859;
      // a box hit counts when tfar >= tmin, :695-700)
      int best = -1;
      float bestT = FLT_MAX;
      for (int i = 0; i < sc.num_volumes; ++i) {
        KRay lr = to_local(sc.vol[i], ray);
        BoxHit bh = intersect_box(lr, sc.vol[i].blo, sc.vol[i].bhi);
        if (!bh.hit || !(bh.tfar >= lr.tmin)) continue;
        float t = fmaxf(lr.tmin, bh.tnear);
        if (t < bestT) {
          bestT = t;
          best = i;
        }
      }
      hit_vol = best;
      // the march over volume `best`, addressed through a wave-uniform index
      // so that its fields are scalar loads
      for (int vi = 0; vi < sc.num_volumes; ++vi) {
        if (vi != best) continue;
        const KVolume &v = sc.vol[__builtin_amdgcn_readfirstlane(vi)];
        KRay lr = to_local(v, ray);
        BoxHit bh = intersect_box(lr, v.blo, v.bhi);
        const float tmin = fmaxf(lr.tmin, bh.tnear);
        const float tmax = fminf(lr.tmax, bh.tfar);
        const f3 ori = v.S * (lr.ori + v.T);
        const f3 dir = v.S * lr.dir;  // not renormalised (VolumeIntegration.h:395-396)
        const float dt = v.cell_size * P.sampling_rate_inv;
        const float expo = dt / v.unit_distance;
        float cr = 0.f, cg = 0.f, cb = 0.f, alpha = 0.f, trans = 1.f;
        // steps needed to cross [tmin, tmax] plus slack; a zero or non-finite
        // step size marches nothing (the reference would not terminate)
        const float span = (tmax - tmin) / dt;
        const int max_steps =
            (dt > 0.f && span < 16777216.f) ? int(fmaxf(span, 0.f) * 1.001f) + 4 : 0;
        int step = 0;
        float t = tmin;
        bool marching = true;
        while (marching) {
          float val[kRaySteps], tb[kRaySteps];
          bool in[kRaySteps], has[kRaySteps];
#pragma unroll
          for (int b = 0; b < kRaySteps; ++b) {
            // the same t sequence as for (t = tmin; ...; t += dt)
            in[b] = t <= tmax && step + b < max_steps;
            if constexpr (ANYF) {
              has[b] = in[b] && sample_field<ANYF>(v, ori + dir * t, val[b]);
            } else {
              // structured fields: every position reads clamped, in-bounds
              // taps, so the batch's fetches go out together, unconditionally
              val[b] = sample_structured(v, ori + dir * t);
              has[b] = in[b];
            }
            tb[b] = t;
            t += dt;
          }
#pragma unroll
          for (int b = 0; b < kRaySteps; ++b) {
            marching = marching && in[b] && alpha < 0.99f;
            if (!marching) continue;
            ++lookups;
            if (!has[b]) continue;  // VolumeIntegration.h:412
            const float4 s = classify(v, s_tf, val[b]);
            const float stepT = pow_ieee(1.f - s.w, expo, s_pow);
            const float k = trans * (1.f - stepT);
            f3 c = mk(s.x, s.y, s.z);
            if constexpr (SHADE) c = gradient_shade<ANYF>(sc, v, P, ori + dir * tb[b], -dir, c, rs);
            cr += k * c.x;
            cg += k * c.y;
            cb += k * c.z;
            alpha += k;
            trans *= stepT;
          }
          step += kRaySteps;
        }
        // over(alpha*float4(color,1), bg) (Raycast_impl.cpp:232)
        const float a = 1.f - alpha;
        result = make_float4(alpha * cr + a * result.x, alpha * cg + a * result.y,
                             alpha * cb + a * result.z, alpha + a * result.w);
      }
    }
    if (gbuf_wanted(F)) {
      // a volume box hit sets the object and instance ids and zeroes normal
      // and albedo; depth and primitive id stay the surface's (or the
      // defaults)
      for (int j = 0; j < sc.num_volumes; ++j)
        if (const int i = uidx(j); i == hit_vol) {
          gb_obj = sc.vol[i].obj_id;
          gb_inst = sc.vol[i].inst_id;
          gb_n = mk(0.f);
          gb_alb = mk(0.f);
        }
      const size_t i = size_t(y) * F.width + x;
      if (gb_depth <= F.depth[i]) gbuf_write(F, i, gb_n, gb_alb, gb_prim, gb_obj, gb_inst);
    }
    write_sample(F, x, y, 0, result, gb_depth, false);
  }
}

// The ray caster.  Persistent: the grid fills the chip once and every wave
// claims 8x8-pixel tiles from a work queue until the image is done, so waves
// of short rays (background, early saturation) hand their SIMD to the next
// tile at once instead of waiting for workgroup dispatch, and the frame has
// no tail of half-empty CUs.  The row-major tile sequence is cut into 8 contiguous
// bands; band q is queue q, drained first by the workgroups the hardware
// placed on XCD q (blockIdx % 8), which then help the other bands.
template <bool ANYF, bool SHADE, bool SURF, bool INST>
__global__ void __launch_bounds__(256)
    k_raycast(KScene sc, KCamera cam, KParams P, KFrame F, uint32_t *queues,
              uint32_t *queues_next) {
  extern __shared__ float4 s_tf[];
  __shared__ float2 s_pow[64];  // pow_ieee's log2 table
  if (threadIdx.x < 64) s_pow[threadIdx.x] = kPowLog2[threadIdx.x];
  if (SURF) stage_nodes(sc, s_tf);
  stage_tf(sc, s_tf);  // (its barrier also publishes s_pow and the nodes)
  // clear the counters the next launch will use (the current ones were
  // cleared by the previous launch)
  if (blockIdx.x == 0 && threadIdx.x < kQueues) queues_next[threadIdx.x] = 0u;
  const int lane = int(threadIdx.x) & 63;
  const int ntx = (F.width + 7) >> 3;
  const uint32_t ntiles = uint32_t(ntx) * uint32_t((F.height + 7) >> 3);
  uint32_t lookups = 0;
  int q = int(blockIdx.x) & (kQueues - 1);
  for (int tries = 0; tries < kQueues;) {
    const uint32_t lo = ntiles * uint32_t(q) / kQueues;
    const uint32_t hi = ntiles * uint32_t(q + 1) / kQueues;
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(&queues[q], 1u);
    k = __shfl(k, 0, 64);
    if (lo + k >= hi) {  // band drained: help the next one
      q = (q + 1) & (kQueues - 1);
      ++tries;
      continue;
    }
    const int tile = int(lo + k);
    const int x = (tile % ntx) * 8 + (lane & 7), y = (tile / ntx) * 8 + (lane >> 3);
    raycast_pixel<ANYF, SHADE, SURF, INST>(sc, cam, P, F, x, y, s_tf, s_pow, lookups);
  }
  if (F.counters) wave_add_counter(&F.counters[0], lookups);
}

// ---------------------------------------------------------------------------
// Path tracer: Pathtrace_impl.cpp:489-641, volume vertices with directional
// light NEE and the ambient-occlusion term, Russian roulette drawing from the
// un-advanced pixel RNG (Pathtrace_impl.cpp:594-599).
//
// Persistent kernel with path regeneration.  The reference runs one thread
// per pixel to completion, so a 64-lane wave lives as long as its longest
// path.  Here every lane owns one path at a time: each loop iteration traces
// one ray (radiance, shadow or AO pass) for every active lane, and a lane
// whose path ended claims the next pixel from its XCD's work queue.  A pixel's
// random streams depend only on (pixelID, accumID), so the image is
// identical to the one-thread-per-pixel schedule.
//
// Work distribution: the image is cut into 32x32-pixel chunks (16 tiles of
// 8x8); chunk c belongs to queue c % 8, and workgroup b drains queue b % 8
// first (the XCD the hardware placed it on), then helps the other queues.
// A wave claims 64 pixel indices per atomic and hands them to its idle
// lanes.
// ---------------------------------------------------------------------------
enum RayType { Radiance = 0, Shadow = 1, AO = 2, Miss = 3 };

struct PathState {
  int pixel;  // linear pixel index, -1 = idle
  uint32_t rng, srng;
  int sampleID;
  float4 acc;
  float depth, sdepth;
  // current ray; tmin is 0 for every ray of a path, and shadow / AO rays
  // start at the vertex's offset position hitPosOff, which the continuation
  // ray of the bounce starts from as well
  f3 ori, dir;
  float tmax;
  int rayType;
  int bounceID;
  bool first;  // the sample's first pass
  bool taa_take;  // TAA: this sample is the pixel's closest so far
  bool firstHit;
  f3 throughput, intensity;
  // ShadeState (Pathtrace_impl.cpp:15-35) of the current vertex
  f3 albedo;       // throughput factor of the BSDF sample: f * (cosTheta / pdf)
  f3 base;         // baseColor (ambient term)
  f3 direct;       // shadedColor * lightVisibility
  float bs_pdf;    // BSDF sample pdf (MIS at a light hit)
  float misw;      // misWeightNEE
  f3 sn, bs_dir;
  int aoSamples;
  float aoWeights, aoCount;
  // radiance cache: query / training slot, training-path flag, radiance and
  // throughput at the second vertex
  int qslot, tslot;
  bool train_sel;
  f3 I2, T2;
};

// Wave-aggregated slot claim: every active lane calls it, lanes with `want`
// receive consecutive indices from one atomic per wave.
__device__ inline uint32_t wave_claim(uint32_t *ctr, bool want) {
  const unsigned long long m = __ballot(want);
  if (m == 0ull) return 0xFFFFFFFFu;
  const uint32_t lane = __lane_id();
  const uint32_t rank = uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (int(lane) == leader) base = atomicAdd(ctr, uint32_t(__popcll(m)));
  base = __shfl(base, leader, 64);
  return want ? base + rank : 0xFFFFFFFFu;
}

// Raw cache inputs of a scattering vertex: position in the cache box,
// direction as (theta/pi, phi/2pi), classified opacity.
__device__ inline void cache_inputs(const KNrc &N, f3 pos, f3 dir, float density, float *raw) {
  const f3 p = (pos - N.pos_lo) * N.pos_scale;
  raw[0] = fminf(fmaxf(p.x, 0.f), 1.f);
  raw[1] = fminf(fmaxf(p.y, 0.f), 1.f);
  raw[2] = fminf(fmaxf(p.z, 0.f), 1.f);
  const f3 d = normalize(dir);
  raw[3] = acosf(fminf(fmaxf(d.z, -1.f), 1.f)) * 0.318309886f;
  raw[4] = (atan2f(d.y, d.x) + 3.14159265f) * 0.159154943f;
  raw[5] = density;
}

constexpr int kChunkLog2 = 5;  // 32x32-pixel chunks

__device__ inline void chunk_pixel(int q, uint32_t local, int nchx, int &x, int &y) {
  const uint32_t j = local >> (2 * kChunkLog2);
  const uint32_t w = local & ((1u << (2 * kChunkLog2)) - 1);
  const uint32_t c = j * kQueues + uint32_t(q);
  const uint32_t tile = w >> 6;
  x = int((c % uint32_t(nchx)) << kChunkLog2) + int((tile & 3) << 3) + int(w & 7);
  y = int((c / uint32_t(nchx)) << kChunkLog2) + int((tile >> 2) << 3) + int((w >> 3) & 7);
}

// Start sample `sampleID` of the lane's pixel: jitter, camera ray, shutter
// draw (Pathtrace_impl.cpp:526-535).  LENS = false compiles the thin lens out
// (the radiance-cache variants, whose entry point refuses a lens camera: the
// lens code there costs spills in the pass loop)
template <bool LENS>
__device__ inline void begin_sample(PathState &st, const KCamera &cam, const KParams &P, int x,
                                    int y, int W, int H) {
  float xf = float(x), yf = float(y);
  const float jx = rnd(st.srng) - .5f;
  const float jy = rnd(st.srng) - .5f;
  xf += jx;
  yf += jy;
  const KRay r = primary_ray(cam, xf, yf, float(W), float(H));
  st.ori = r.ori;
  st.dir = r.dir;
  st.tmax = r.tmax;
  if (LENS && cam.lens_u > 0.f) {  // the lens sample precedes the shutter draw
    const float u1 = rnd(st.srng);
    const float u2 = rnd(st.srng);
    thin_lens(st.ori, st.dir, cam, u1, u2);
  }
  (void)rnd(st.srng);  // shutter time (DeviceCopyableObjects.h:2597)
  st.sdepth = 1e31f;
  st.rayType = Radiance;
  st.bounceID = 0;
  st.first = true;
  st.firstHit = false;
  st.throughput = mk(1.f);
  st.intensity = mk(0.f);
  st.bs_pdf = 0.f;
  st.qslot = -1;
  st.tslot = -1;
}

// intersectAll (renderer/common.h:1606-1620) of one pass: surfaces (closest
// hit; any hit for shadow / AO rays), quad lights (invisible ones skipped for
// camera rays), volumes (delta tracking, closest); a light wins ties with a
// surface or volume, a volume must be strictly closer than a surface.
enum HitKind { kNone = 0, kVolume = 1, kSurface = 2, kLight = 3 };

struct PassHit {
  int kind;
  bool surf_hit;     // any surface hit in [tmin, tmax], whichever kind wins
  float trans;       // ratio tracking: the volumes' transmittance estimate (else 1)
  float t;
  f3 albedo;         // volume: classified color
  float extinction;  // volume: classified opacity
  int id;            // surface, light or volume index
  uint32_t rec;      // surface: triangle record; volume: sample_field's record
};

// AREA: the scene has surfaces or quad lights (else the kernel is built
// without the BVH and quad code, which would only add registers); ANYF: it
// has fields other than structuredRegular
template <bool AREA, bool ANYF, bool INST>
__device__ inline PassHit intersect_all(const KScene &sc, const KRay &ray, bool shadow,
                                        int bounceID, const float4 *s_tf, const float *steps,
                                        uint32_t &srng, uint32_t &lookups, int2 *bstack,
                                        const float4 *lnodes, bool ratio = false) {
  bool sh = false;
  float st = ray.tmax;
  int sid = -1;
  uint32_t srec = 0;
  for (int i = 0; AREA && i < sc.num_surfaces; ++i) {
    float t = st;
    int prim = -1;
    uint32_t rec = 0;
    if (bvh_hit(sc.surf[i], ray, shadow, t, prim, rec, bstack, lnodes + sc.staged[i].lds_first,
                sc.staged[i].lds_nodes)) {
      sh = true;
      st = t;
      sid = i;
      srec = rec;
    }
  }
  if (AREA && INST && sc.tls.num_inst > 0 && !(shadow && sh)) {
    float t = st;
    int slot = -1, prim = -1;
    uint32_t rec = 0;
    if (tls_hit(sc.tls, ray, shadow, sh, t, slot, prim, rec, bstack)) {
      sh = true;
      st = t;
      sid = sc.num_surfaces + slot;
      srec = rec;
    }
  }
  bool lh = false;
  float lt = FLT_MAX;
  int lid = -1;
  for (int i = 0; AREA && i < sc.num_lights; ++i) {
    const KLight &L = sc.light[i];
    if (bounceID == 0 && !L.visible) continue;
    if (L.type == VRT_LIGHT_HDRI) {
      // every ray meets the environment, at FLT_MAX (common.h:1574-1588)
      if (!lh) {
        lh = true;
        lt = FLT_MAX;
        lid = i;
      }
      continue;
    }
    if (L.type != VRT_LIGHT_QUAD) continue;
    float t;
    if (quad_hit(L, ray, t) && t < lt) {
      lh = true;
      lt = t;
      lid = i;
    }
  }
  VolHit vh;
  vh.hit = false;
  vh.t = FLT_MAX;
  int vid = -1;
  float trans = 1.f;
  for (int i = 0; i < sc.num_volumes; ++i) {
    if (ratio) {  // shadow / AO ray: the product of the volumes' estimates
      trans = trans * ratio_track<ANYF>(sc.vol[i], s_tf, lds_majorants(sc, s_tf, i), steps[i],
                                        to_local(sc.vol[i], ray), srng, lookups);
      continue;
    }
    VolHit h = delta_track<ANYF>(sc.vol[i], s_tf, lds_majorants(sc, s_tf, i), steps[i],
                                 to_local(sc.vol[i], ray), srng, lookups);
    if (h.hit && h.t < vh.t) {
      vh = h;
      vid = i;
    }
  }
  PassHit out;
  out.kind = kNone;
  out.surf_hit = sh;
  out.trans = trans;
  out.t = FLT_MAX;
  out.id = -1;
  out.rec = 0;
  if (lh && (!sh || lt <= st) && (!vh.hit || lt <= vh.t)) {
    out.kind = kLight;
    out.t = lt;
    out.id = lid;
  } else if (vh.hit && (!sh || vh.t < st) && (!lh || vh.t < lt)) {
    out.kind = kVolume;
    out.t = vh.t;
    out.albedo = vh.albedo;
    out.extinction = vh.extinction;
    out.id = vid;
    out.rec = vh.rec;
  } else if (sh) {
    out.kind = kSurface;
    out.t = st;
    out.id = sid;
    out.rec = srec;
  } else if (vh.hit || lh) {
    // a hit that is neither (e.g. a volume behind a light): hitRec.hit holds
    out.kind = vh.hit ? kVolume : kLight;
    out.t = vh.hit ? vh.t : lt;
    out.albedo = vh.albedo;
    out.extinction = vh.extinction;
    out.id = vh.hit ? vid : lid;
    out.rec = vh.hit ? vh.rec : 0u;
  }
  return out;
}


// Russian roulette, run after every pass that continues the path once
// bounceID > 1 -- shadow and AO passes included -- drawing from the
// un-advanced pixel RNG (Pathtrace_impl.cpp:594-599).  True: the path ends.
__device__ inline bool roulette(PathState &st) {
  if (st.bounceID > 1) {
    const float prob = max3(st.throughput);
    if (rnd(st.rng) > prob) return true;
    st.throughput = st.throughput * (1.f / prob);
  }
  return false;
}

// project (visionaray, parity unpinned: gluProject) of a world point with a
// row-major proj * view matrix into window coordinates of a W x H viewport
__device__ inline float2 project_window(const float *M, f3 p, int W, int H) {
  const float x = ((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3];
  const float y = ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7];
  const float w = ((M[12] * p.x + M[13] * p.y) + M[14] * p.z) + M[15];
  return make_float2(float(W) * (x / w + 1.f) / 2.f, float(H) * (y / w + 1.f) / 2.f);
}

// TAA at a radiance vertex of the pixel's closest sample: the motion vector
// prev - curr window position of the hit, written at every such vertex
// (the last one stands, Pathtrace_impl.cpp:290-297), and the first hit's
// albedo
__device__ inline void taa_vertex(const KFrame &F, const KParams &P, const PathState &st,
                                  f3 hitPos, f3 albedo) {
  const float2 a = project_window(P.vp_prev, hitPos, F.width, F.height);
  const float2 b = project_window(P.vp_curr, hitPos, F.width, F.height);
  F.motion[st.pixel] = make_float4(a.x - b.x, a.y - b.y, 0.f, 1.f);
  if (st.bounceID == 0) {
    float *al = F.taa_curr_alb + 3 * size_t(st.pixel);
    al[0] = albedo.x;
    al[1] = albedo.y;
    al[2] = albedo.z;
  }
}

// One pass of the bounce loop (Pathtrace_impl.cpp:558-600) after the ray's
// intersection `h`: shade, advance the state machine, set up the next ray.
// Returns true when the sample's path ended.  With NRC, a path reaching its
// query vertex (the second volume vertex, or the first) either becomes a
// cache query and ends, or -- a training path -- records the vertex and goes
// on.
template <bool NRC, bool AREA, bool ANYF, bool GRAD>
__device__ inline bool path_shade(PathState &st, const KScene &sc, const KParams &P,
                                  const KNrc &N, const KFrame &F, const PassHit &h) {
  KRay ray;
  ray.ori = st.ori;
  ray.dir = st.dir;
  const bool hit = h.kind != kNone;
  if (st.first) {
    st.first = false;
    st.firstHit = hit;
    if (h.kind == kVolume || h.kind == kSurface) st.sdepth = h.t;
    // TAA keeps the closest sample's albedo and motion (the first of equal
    // depths, Pathtrace_impl.cpp:616)
    st.taa_take = P.taa && (h.kind == kVolume || h.kind == kSurface) && h.t < st.depth;
  }
  if constexpr (NRC) {
    const bool v2 = st.rayType == Radiance && h.kind == kVolume && st.bounceID == N.qbounce;
    (void)wave_claim(&N.ctr[2], v2);  // vertex-2 paths of this frame
    const bool want_train = v2 && st.train_sel;
    const uint32_t tslot = wave_claim(&N.ctr[1], want_train);
    const bool is_train = want_train && tslot < uint32_t(N.t_cap);
    const bool want_query = v2 && !is_train;
    const uint32_t qslot = wave_claim(&N.ctr[0], want_query);
    // self-training: a training path whose vertex textra bounces after its
    // query vertex is a volume vertex ends there in a cache query of its own;
    // the query's prediction completes both the pixel (through the resolve)
    // and the training target
    const bool term = st.tslot >= 0 && N.textra > 0 && st.rayType == Radiance &&
                      h.kind == kVolume && st.bounceID == N.qbounce + N.textra;
    const uint32_t tq = wave_claim(&N.ctr[0], term);
    if (term) {
      if (tq < uint32_t(N.q_cap)) {
        float raw[6];
        cache_inputs(N, ray.ori + ray.dir * h.t, ray.dir, h.extinction, raw);
        for (int i = 0; i < 6; ++i) N.q_raw[6 * size_t(tq) + i] = raw[i];
        N.q_T[3 * size_t(tq)] = st.throughput.x;
        N.q_T[3 * size_t(tq) + 1] = st.throughput.y;
        N.q_T[3 * size_t(tq) + 2] = st.throughput.z;
        st.qslot = int(tq);
        N.t_query[st.tslot] = int(tq);
        float *r = N.t_ratio + 3 * size_t(st.tslot);
        r[0] = st.T2.x > 0.f ? st.throughput.x / st.T2.x : 0.f;
        r[1] = st.T2.y > 0.f ? st.throughput.y / st.T2.y : 0.f;
        r[2] = st.T2.z > 0.f ? st.throughput.z / st.T2.z : 0.f;
      }
      return true;
    }
    if (v2) {
      float raw[6];
      cache_inputs(N, ray.ori + ray.dir * h.t, ray.dir, h.extinction, raw);
      if (is_train) {
        for (int i = 0; i < 6; ++i) N.t_raw[6 * size_t(tslot) + i] = raw[i];
        st.tslot = int(tslot);
        st.I2 = st.intensity;
        st.T2 = st.throughput;
        if (N.textra > 0) N.t_query[tslot] = -1;
      } else {
        if (qslot < uint32_t(N.q_cap)) {
          for (int i = 0; i < 6; ++i) N.q_raw[6 * size_t(qslot) + i] = raw[i];
          N.q_T[3 * size_t(qslot)] = st.throughput.x;
          N.q_T[3 * size_t(qslot) + 1] = st.throughput.y;
          N.q_T[3 * size_t(qslot) + 2] = st.throughput.z;
          st.qslot = int(qslot);
        }
        return true;  // the cache supplies the radiance from here on
      }
    }
  }
  // shade (Pathtrace_impl.cpp:93-487)
  const int rayType = st.rayType;
  if (rayType == Radiance) {
    // a radiance miss ends the path; its shade state adds nothing
    if (!hit) return true;
    const int nl = sc.num_lights;
    if (AREA && h.kind == kLight) {
      // the quad's intensity, MIS-weighted against the BSDF sample after the
      // first bounce (Pathtrace_impl.cpp:139-183); the path ends
      KLight L = sc.light[0];
      for (int j = 1; j < nl; ++j)
        if (const int i = uidx(j); i == h.id) L = sc.light[i];
      if (L.type == VRT_LIGHT_HDRI) {
        // the map in the ray's direction, MIS-weighted after the first
        // bounce (Pathtrace_impl.cpp:144-182)
        f3 shaded = hdri_radiance(sc.hdri, ray.dir);
        if (st.bounceID > 0)
          shaded = shaded * power_heuristic(st.bs_pdf, hdri_pdf(sc.hdri, ray.dir) / float(nl));
        st.intensity = st.intensity + st.throughput * ((shaded * 1.f) * 1.f);
        st.intensity = st.intensity + st.throughput * (mk(0.f) * P.ambient_color *
                                                       P.ambient_radiance * 1.f);
        return true;
      }
      f3 shaded = L.intensity;
      if (st.bounceID > 0) {
        const f3 hitPos = ray.ori + ray.dir * h.t;
        const float A = L.a1 + L.a2;
        const float ld = length(hitPos - ray.ori);
        const f3 Ld = normalize(hitPos - ray.ori);
        const float LdotNl = fabsf(dot(-Ld, L.nrm));
        const float solidAngle = (LdotNl * A) / (ld * ld);
        const float lightPDF = 1.f / solidAngle;
        shaded = shaded * power_heuristic(st.bs_pdf, lightPDF / float(nl));
      }
      const f3 ambient = mk(0.f) * P.ambient_color * P.ambient_radiance * 1.f;
      st.intensity = st.intensity + st.throughput * ((shaded * 1.f) * 1.f);
      st.intensity = st.intensity + st.throughput * ambient;
      return true;
    }
    const f3 viewDir = -normalize(ray.dir);
    const f3 hitPos = ray.ori + ray.dir * h.t;
    const float eps = epsilon_from(hitPos, ray.dir, h.t);
    f3 gn = mk(0.f), color;
    const bool gbuf = st.bounceID == 0 && gbuf_wanted(F) && h.t < st.depth &&
                      h.t <= F.depth[st.pixel];
    if (h.kind == kVolume) {
      if constexpr (GRAD) {
        // the gradient at the hit in voxel space (Pathtrace_impl.cpp:204-212),
        // the lanes of each volume in turn
        for (int j = 0; j < sc.num_volumes; ++j) {
          if (uidx(j) != h.id) continue;
          const KVolume &hv = sc.vol[opaque_uidx(uidx(j))];
          const KRay lr = to_local(hv, ray);
          gn = gradient_normal<ANYF>(hv, hv.S * ((lr.ori + lr.dir * h.t) + hv.T));
        }
      }
      if (P.ambient_samples > 0 && length(gn) < 1e-3f) {
        const float u1 = rnd(st.srng), u2 = rnd(st.srng);
        gn = uniform_sample_sphere(u1, u2);
      }
      if (GRAD || P.ambient_samples > 0) {
        // normal transform transpose(affineInv) (Pathtrace_impl.cpp:217-219,
        // Instance.cpp:96-98), as mat3 * vec3 = col0 x + col1 y + col2 z;
        // the volume's matrix picked by a select loop (a divergent index into
        // the kernel argument would copy it to scratch)
        float A[9];
        for (int k = 0; k < 9; ++k) A[k] = sc.vol[0].A[k];
        for (int j = 1; j < sc.num_volumes; ++j) {
          const int i = uidx(j);
          const bool m = i == h.id;
          for (int k = 0; k < 9; ++k) A[k] = m ? sc.vol[i].A[k] : A[k];
        }
        gn = mk(A[0] * gn.x + A[3] * gn.y + A[6] * gn.z, A[1] * gn.x + A[4] * gn.y + A[7] * gn.z,
                A[2] * gn.x + A[5] * gn.y + A[8] * gn.z);
      }
      st.sn = faceforward(gn, viewDir, gn);
      gn = faceforward(gn, viewDir, gn);
      color = h.albedo;
      if (gbuf) {
        uint32_t obj = 0, inst = 0, prim = 0;
        for (int j = 0; j < sc.num_volumes; ++j)
          if (const int i = uidx(j); i == h.id) {  // (ids only: cheap even if folded)
            obj = sc.vol[i].obj_id;
            inst = sc.vol[i].inst_id;
            if (h.rec != kNoRec) prim = sc.vol[i].ug_cell[h.rec];
          }
        gbuf_write(F, size_t(st.pixel), gn, color, prim, obj, inst);
      }
      if (st.taa_take) taa_vertex(F, P, st, hitPos, color);
    } else if (AREA) {
      // triangle: Ng = normalize(e1 x e2), no vertex normals (getNormals,
      // common.h:233-264), normalised again after the instance's normal
      // transform (Pathtrace_impl.cpp:243-247); matte color
      const SurfRef R = surface_ref(sc, h.id);
      const float4 v1 = R.tris[3 * size_t(h.rec)];
      gn = surface_normal(R, h.rec);
      st.sn = faceforward(gn, viewDir, gn);
      gn = faceforward(gn, viewDir, gn);
      color = R.color;
      if (gbuf)
        gbuf_write(F, size_t(st.pixel), gn, color, __float_as_uint(v1.w), R.obj_id, R.inst_id);
      if (st.taa_take) taa_vertex(F, P, st, hitPos, color);
    }
    LightSmp ls;
    ls.intensity = mk(0.f);
    ls.dir = mk(0.f);
    ls.pdf = ls.dist = ls.dist2 = 0.f;
    ls.type = -1;
    if (nl > 0) {
      int which = int(rnd(st.srng) * float(nl));
      if (which == nl) which = 0;
      KLight L = sc.light[0];
      for (int j = 1; j < nl; ++j)
        if (const int i = uidx(j); i == which) L = sc.light[i];
      ls = sample_light<AREA>(sc, L, hitPos, st.srng);
    }
    const f3 lightIntensity = ls.intensity * safe_rcp(ls.dist2);
    if (h.kind == kVolume) {
      // Pathtrace_impl.cpp:330-339: no phase factor, no light count;
      // isotropic phase sample: f = albedo, pdf = cosTheta = 1.  With an
      // anisotropic phase (g != 0) NEE carries 4 pi p_HG towards the light
      // and the sample is drawn from p_HG, whose f/pdf is again the albedo.
      // gradient shading: matte NEE along the gradient normal (:311-328)
      const bool matte = GRAD && length(gn) > 1e-10f;
      if (matte) {
        const float inv_pi = 0.318309886183790671538f;
        const float NdotL = fmaxf(0.f, dot(st.sn, normalize(ls.dir)));
        st.direct = h.albedo * inv_pi * lightIntensity * NdotL * safe_rcp(ls.pdf) * float(nl);
      } else {
        st.direct = h.albedo * lightIntensity * safe_rcp(ls.pdf) * safe_rcp(ls.dist2);
      }
      const float u1 = rnd(st.srng), u2 = rnd(st.srng);
      if (P.phase_g == 0.f) {
        st.bs_dir = uniform_sample_sphere(u1, u2);
        st.bs_pdf = 1.f;
      } else {
        const f3 d = normalize(ray.dir);
        if (nl > 0 && !matte) st.direct = st.direct * hg_4pi(P.phase_g, dot(d, normalize(ls.dir)));
        st.bs_dir = hg_sample(P.phase_g, d, u1, u2);
        st.bs_pdf = hg_4pi(P.phase_g, dot(d, st.bs_dir));
      }
      st.albedo = h.albedo * (1.f * 1.f);
    } else if (AREA) {
      // matte (common.h:1033-1048, 1371-1413): f = color / pi, cosine
      // hemisphere around faceforward(Ns, viewDir, Ng), pdf = cos / pi
      const float inv_pi = 0.318309886183790671538f;
      const f3 f = color * inv_pi;
      const f3 lightDir = normalize(ls.dir);
      const float NdotL = fmaxf(0.f, dot(st.sn, lightDir));
      st.direct = f * lightIntensity * NdotL * safe_rcp(ls.pdf) * float(nl);
      const f3 w = faceforward(st.sn, viewDir, gn);
      const f3 v = fabsf(w.x) > fabsf(w.y) ? normalize(mk(-w.z, 0.f, w.x))
                                           : normalize(mk(0.f, w.z, -w.y));
      const f3 u = cross(v, w);
      const float u1 = rnd(st.srng), u2 = rnd(st.srng);
      const f3 sp = cosine_sample_hemisphere(u1, u2);
      st.bs_dir = normalize(u * sp.x + v * sp.y + w * sp.z);
      st.bs_pdf = fmaxf(0.f, dot(st.sn, st.bs_dir)) * inv_pi;
      const float cosT = fmaxf(0.f, dot(st.sn, st.bs_dir));
      st.albedo = f * (cosT * safe_rcp(st.bs_pdf));
    }
    // misWeightNEE (Pathtrace_impl.cpp:370-380)
    st.misw = nl > 0 ? ((ls.type == VRT_LIGHT_QUAD || ls.type == VRT_LIGHT_HDRI)
                            ? power_heuristic(ls.pdf / float(nl), st.bs_pdf)
                            : 1.f)
                     : 0.f;
    st.base = color;
    st.aoSamples = 0;
    st.aoWeights = 0.f;
    st.aoCount = 0.f;
    // lightSample.pdf >= 0 always holds: shadow ray next (:403-406)
    st.ori = hitPos + st.sn * eps;  // hitPosOff
    st.dir = normalize(ls.dir);
    st.tmax = nl > 0 ? ls.dist : 0.f;
    st.rayType = Shadow;
    return roulette(st);
  }
  float vis_ao = 1.f;
  if (rayType == Shadow) {
    // visibility = surfV * volV * Tr (Pathtrace_impl.cpp:439-444): blocked
    // by any surface hit or by a volume collision that wins the pass; with
    // ratio tracking Tr is the volumes' transmittance estimate
    st.direct = st.direct * ((h.surf_hit || h.kind == kVolume) ? 0.f : h.trans);
  } else {  // AO
    st.aoSamples++;
    const float weight = fmaxf(0.f, dot(ray.dir, st.sn));
    st.aoWeights += weight;
    // occluded by any hit; with ratio tracking the volumes occlude by 1 - Tr
    if (weight > 0.f) st.aoCount += weight * (hit ? 1.f : 1.f - h.trans);
    if (st.aoSamples >= P.ambient_samples)
      vis_ao = st.aoWeights > 0.f ? 1.f - st.aoCount / st.aoWeights : 0.f;
  }
  // after the shadow pass (or an AO pass with samples left) an AO ray is
  // traced while ambient samples remain (Pathtrace_impl.cpp:437-455); the
  // origin stays hitPosOff
  if (st.aoSamples < P.ambient_samples) {
    const f3 w = st.sn;
    const f3 v = fabsf(w.x) > fabsf(w.y) ? normalize(mk(-w.z, 0.f, w.x))
                                         : normalize(mk(0.f, w.z, -w.y));
    const f3 u = cross(v, w);
    const float u1 = rnd(st.srng), u2 = rnd(st.srng);
    const f3 sp = cosine_sample_hemisphere(u1, u2);
    st.dir = normalize(u * sp.x + v * sp.y + w * sp.z);
    st.tmax = P.occlusion_distance;
    st.rayType = AO;
    return roulette(st);
  }
  // the bounce ends: accumulate, continue along the BSDF sample
  const f3 ambient = st.base * P.ambient_color * P.ambient_radiance * vis_ao;
  st.intensity = st.intensity + st.throughput * (st.direct * st.misw);
  st.intensity = st.intensity + st.throughput * ambient;
  st.throughput = st.throughput * st.albedo;
  st.bounceID++;
  if (st.bounceID > P.max_bounce) return true;
  if (roulette(st)) return true;
  st.dir = normalize(st.bs_dir);
  st.tmax = 1e31f;
  st.rayType = Radiance;
  return false;
}

// training-path selection: an independent hash of (pixel, accumID), so that
// the path's own random streams are the ones of the plain path tracer
__device__ inline float train_hash(uint32_t pixel, uint32_t accumID) {
  return float(tea4(pixel ^ 0x5bd1e995u, accumID + 0x68e31da4u) & 0x00FFFFFFu) *
         (1.0f / 16777216.0f);
}

// GRAD: gradient shading (the renderer's gradientShading), compiled only
// into the variants that use it
// 3 waves per SIMD (<= 168 VGPRs): the variants with surfaces or other
// field types would otherwise settle at 2 and lose a third of their latency
// hiding; the compiler fits them without spilling
template <bool NRC, bool AREA, bool ANYF, bool GRAD, bool INST>
__global__ void __launch_bounds__(256, 3)
    k_pathtrace(KScene sc, KCamera cam, KParams P, KFrame F, int accumID, uint32_t *queues,
                uint32_t *queues_next, KNrc N) {
  extern __shared__ float4 s_tf[];
  if (AREA) stage_nodes(sc, s_tf);
  stage_tf(sc, s_tf);  // (its barrier also publishes the nodes)
  // clear the counters the next launch will use (the current ones were
  // cleared by the previous launch)
  if (blockIdx.x == 0 && threadIdx.x < kQueues) queues_next[threadIdx.x] = 0u;
  // fraction of paths that become training paths: the per-frame training
  // budget over the previous frame's count of paths reaching vertex 2
  float p_train = 0.f;
  if constexpr (NRC) {
    if (N.train) {
      const uint32_t prev = N.ctr[3];
      p_train = prev == 0u ? 1.f : fminf(1.f, float(N.t_cap) / float(prev));
    }
  }

  const int W = F.width, H = F.height;
  const int nchx = (W + (1 << kChunkLog2) - 1) >> kChunkLog2;
  const int nchy = (H + (1 << kChunkLog2) - 1) >> kChunkLog2;
  const uint32_t nch = uint32_t(nchx * nchy);
  const uint32_t lane = threadIdx.x & 63u;
  const unsigned long long below = (1ull << lane) - 1ull;
  const float *steps = lds_steps(sc, s_tf);
  // this lane's BVH traversal stack, after the tables and majorants
  int2 *bstack = AREA ? reinterpret_cast<int2 *>(reinterpret_cast<char *>(s_tf) +
                                                 scene_lds_bytes(sc)) +
                            threadIdx.x
                      : nullptr;
  const float4 *lnodes = AREA ? staged_nodes(sc, s_tf) : nullptr;

  PathState st;
  st.pixel = -1;
  uint32_t lookups = 0, hit_samples = 0;
  int q = int(blockIdx.x) & (kQueues - 1);  // wave-uniform queue
  int qtries = 0;
  uint32_t pool_base = 0, pool_left = 0;    // wave-uniform local pool
  int px = 0, py = 0;

  // chain: this lane's bounce continues with its shadow / AO rays; while any
  // lane of the wave has such a ray (P.chain_passes), the pass is theirs and
  // nothing is refilled
  bool chain = false;
  while (true) {
    const bool chained = __ballot(chain) != 0ull;
    // ---- refill idle lanes ----------------------------------------------------
    unsigned long long idle = chained ? 0ull : __ballot(st.pixel < 0);
    if (P.refill_min > 1 && __popcll(idle) < P.refill_min && __ballot(st.pixel >= 0) != 0ull)
      idle = 0ull;  // refill later, for more lanes at once
    while (idle && qtries < kQueues) {
      if (pool_left == 0) {
        // pixels of queue q: the chunks c = q, q+8, ... times 1024 pixels
        const uint32_t band = ((nch + kQueues - 1 - uint32_t(q)) / kQueues) << (2 * kChunkLog2);
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&queues[q], 64u);
        base = __shfl(base, 0, 64);
        if (base >= band) {
          q = (q + 1) & (kQueues - 1);
          ++qtries;
          continue;
        }
        pool_base = base;
        pool_left = min(64u, band - base);
      }
      const uint32_t nidle = uint32_t(__popcll(idle));
      const uint32_t take = min(nidle, pool_left);
      const uint32_t rank = uint32_t(__popcll(idle & below));
      chunk_pixel(q, pool_base + rank, nchx, px, py);
      if (st.pixel < 0 && rank < take && px < W && py < H) {
        st.pixel = py * W + px;
        st.rng = tea4(uint32_t(st.pixel), uint32_t(accumID) + P.rng_offset);
        st.srng = st.rng;  // ScreenSample's copy (Pathtrace_impl.cpp:512-513)
        st.sampleID = 0;
        st.acc = make_float4(0.f, 0.f, 0.f, 0.f);
        st.depth = 1e31f;
        st.train_sel =
            NRC && train_hash(uint32_t(st.pixel), uint32_t(accumID) + P.rng_offset) < p_train;
        begin_sample<!NRC>(st, cam, P, px, py, W, H);
      }
      pool_base += take;
      pool_left -= take;
      idle = __ballot(st.pixel < 0);
      if (take == nidle) break;  // every idle lane was offered a pixel
    }
    if (!chained && __ballot(st.pixel >= 0) == 0ull) {
      // no lane holds a pixel: either every queue is drained, or the last
      // pool only covered padding pixels outside the image -- claim again
      if (qtries >= kQueues) break;
      continue;
    }

    // ---- one pass for every active lane (or every chained one) -----------------
    if (st.pixel >= 0 && (!chained || chain)) {
      chain = false;
      KRay ray;
      ray.ori = st.ori;
      ray.dir = st.dir;
      ray.tmin = 0.f;
      ray.tmax = st.tmax;
      clip_ray(ray, P);  // every pass (Pathtrace_impl.cpp:559)
      const bool shadow = st.rayType != Radiance;
      const PassHit hit = intersect_all<AREA, ANYF, INST>(sc, ray, shadow, st.bounceID, s_tf, steps,
                                                    st.srng, lookups, bstack, lnodes,
                                                    shadow && P.ratio_tracking != 0);
      if (path_shade<NRC, AREA, ANYF, GRAD>(st, sc, P, N, F, hit)) {
        float4 color = P.bg;
        if (st.firstHit) color = make_float4(st.intensity.x, st.intensity.y, st.intensity.z, 1.f);
        hit_samples += st.firstHit ? 1u : 0u;
        st.acc.x += color.x;
        st.acc.y += color.y;
        st.acc.z += color.z;
        st.acc.w += color.w;
        if (st.sdepth < st.depth) st.depth = st.sdepth;
        const int x = st.pixel % W, y = st.pixel / W;
        if constexpr (NRC) {
          if (st.tslot >= 0) {
            // training target: radiance the path gathered from vertex 2 on,
            // divided by the throughput it arrived there with
            const f3 s = st.intensity - st.I2;
            float *t = N.t_target + 3 * size_t(st.tslot);
            t[0] = st.T2.x > 0.f ? s.x / st.T2.x : 0.f;
            t[1] = st.T2.y > 0.f ? s.y / st.T2.y : 0.f;
            t[2] = st.T2.z > 0.f ? s.z / st.T2.z : 0.f;
          }
          // one sample per pixel: the resolve pass adds the cache term
          N.sample[st.pixel] = st.acc;
          N.query_of_pixel[st.pixel] = st.qslot;
          if (gbuf_wanted(F) && !(st.depth < 1e31f)) gbuf_miss(F, size_t(st.pixel));
          if (F.depth && st.depth <= F.depth[st.pixel]) F.depth[st.pixel] = st.depth;
          st.pixel = -1;
        } else if (++st.sampleID < P.pixel_samples) {
          begin_sample<!NRC>(st, cam, P, x, y, W, H);
        } else {
          const float inv = 1.f / float(P.pixel_samples);
          const float4 a = make_float4(st.acc.x * inv, st.acc.y * inv, st.acc.z * inv,
                                       st.acc.w * inv);
          if (gbuf_wanted(F) && !(st.depth < 1e31f)) gbuf_miss(F, size_t(st.pixel));
          if (P.taa) {
            // TAA: record the sample for the resolve pass (fillGBuffer only,
            // Pathtrace_impl.cpp:636-637); a pixel without hits keeps zero
            // motion and albedo
            const size_t i = size_t(st.pixel);
            if (!(st.depth < 1e31f)) {
              F.motion[i] = make_float4(0.f, 0.f, 0.f, 1.f);
              F.taa_curr_alb[3 * i] = F.taa_curr_alb[3 * i + 1] = F.taa_curr_alb[3 * i + 2] = 0.f;
            }
            F.taa_curr[i] = a;
            if (F.depth && st.depth <= F.depth[i]) F.depth[i] = st.depth;
          } else {
            write_sample(F, x, y, accumID, a, st.depth, true);
          }
          st.pixel = -1;
        }
      } else {
        chain = P.chain_passes && st.rayType != Radiance;
      }
    }
  }
  if (F.counters) {
    wave_add_counter(&F.counters[0], lookups);
    wave_add_counter(&F.counters[1], hit_samples);
  }
}

// TAA resolve (Frame.cpp:339-381, DeviceCopyableObjects.h:2657-2712): the
// history (bilinear over the previous colours, clamp) at the motion-
// compensated position, blended with the frame's sample by alpha -- 1 where
// the albedo at the previous pixel differs by 1e-2 or more -- then tone
// mapped
__global__ void __launch_bounds__(256) k_taa_resolve(KFrame F, float alpha) {
  const int n = F.width * F.height;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = i % F.width, y = i / F.width;
  const float4 mv = F.motion[i];
  int px = int(float(x) + mv.x), py = int(float(y) + mv.y);
  px = min(max(px, 0), F.width - 1);
  py = min(max(py, 0), F.height - 1);
  const size_t pi = size_t(py) * F.width + px;
  float a = alpha;
  const float *pa = F.taa_prev_alb + 3 * pi, *ca = F.taa_curr_alb + 3 * size_t(i);
  if (!(fabsf(pa[0] - ca[0]) < 1e-2f && fabsf(pa[1] - ca[1]) < 1e-2f &&
        fabsf(pa[2] - ca[2]) < 1e-2f))
    a = 1.f;
  // tex2D at ((prevX + 0.5) / W, (prevY + 0.5) / H): texel coordinates prevX, prevY
  const float tx = float(x) + mv.x, ty = float(y) + mv.y;
  const float cx = fminf(fmaxf(tx, -1.f), float(F.width)), cy = fminf(fmaxf(ty, -1.f), float(F.height));
  const float fx = floorf(cx), fy = floorf(cy);
  const float wx = cx - fx, wy = cy - fy;
  const int x0 = min(max(int(fx), 0), F.width - 1), x1 = min(max(int(fx) + 1, 0), F.width - 1);
  const int y0 = min(max(int(fy), 0), F.height - 1), y1 = min(max(int(fy) + 1, 0), F.height - 1);
  const float4 t00 = F.taa_prev[size_t(y0) * F.width + x0], t10 = F.taa_prev[size_t(y0) * F.width + x1];
  const float4 t01 = F.taa_prev[size_t(y1) * F.width + x0], t11 = F.taa_prev[size_t(y1) * F.width + x1];
  auto lp = [](float p, float q, float t) { return p + t * (q - p); };
  const float4 h = make_float4(lp(lp(t00.x, t10.x, wx), lp(t01.x, t11.x, wx), wy),
                               lp(lp(t00.y, t10.y, wx), lp(t01.y, t11.y, wx), wy),
                               lp(lp(t00.z, t10.z, wx), lp(t01.z, t11.z, wx), wy),
                               lp(lp(t00.w, t10.w, wx), lp(t01.w, t11.w, wx), wy));
  const float4 s = F.taa_curr[i];
  const float b = 1.f - a;
  const float4 c = make_float4(b * h.x + a * s.x, b * h.y + a * s.y, b * h.z + a * s.z,
                               b * h.w + a * s.w);
  F.taa_curr[i] = c;
  tone_map(F, size_t(i), c);
}

// Radiance-cache frame: start (counters) and resolve (cache term + frame
// accumulation, Frame::writeSample)
// relMSE = mean over pixels and RGB of (x - r)^2 / (r^2 + eps), x the float4
// running mean, r = ref_sum * ref_scale (FP64 sums); one double atomic per
// workgroup
__global__ void __launch_bounds__(256)
    k_relmse(const float4 *__restrict__ x, const double *__restrict__ ref_sum, double ref_scale,
             int n, float eps, double *out) {
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 v = x[i];
    const double *r = ref_sum + 4 * size_t(i);
    const double r0 = r[0] * ref_scale, r1 = r[1] * ref_scale, r2 = r[2] * ref_scale;
    s += (double(v.x) - r0) * (double(v.x) - r0) / (r0 * r0 + eps);
    s += (double(v.y) - r1) * (double(v.y) - r1) / (r1 * r1 + eps);
    s += (double(v.z) - r2) * (double(v.z) - r2) / (r2 * r2 + eps);
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (int(threadIdx.x) < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, sh[0]);
}

hipError_t launch_relmse(const float4 *x, const double *ref_sum, double ref_scale, int n,
                         float eps, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_relmse, dim3(std::max(1, std::min((n + 255) / 256, 1024))), dim3(256), 0,
                     s, x, ref_sum, ref_scale, n, eps, out);
  return hipGetLastError();
}

__global__ void k_nrc_frame_begin(KNrc N) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    N.ctr[3] = N.ctr[2];
    N.ctr[0] = 0u;
    N.ctr[1] = 0u;
    N.ctr[2] = 0u;
  }
}

// Self-training targets: a training sample whose path ended in a cache
// query adds that query's prediction, scaled by the path's throughput there
// over its throughput at the query vertex, to its target (the suffix radiance
// the path gathered on the way)
__global__ void k_nrc_train_targets(KNrc N) {
  const int n = min(N.t_cap, int(N.ctr[1]));
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int q = N.t_query[i];
    if (q < 0) continue;
    const float *o = N.q_out + 3 * size_t(q);
    const float *r = N.t_ratio + 3 * size_t(i);
    float *t = N.t_target + 3 * size_t(i);
    t[0] += r[0] * fmaxf(o[0], 0.f);
    t[1] += r[1] * fmaxf(o[1], 0.f);
    t[2] += r[2] * fmaxf(o[2], 0.f);
  }
}

__global__ void k_nrc_resolve(KFrame F, KNrc N, int accumID) {
  const int n = F.width * F.height;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 c = N.sample[i];
    const int q = N.query_of_pixel[i];
    if (q >= 0) {
      const float *o = N.q_out + 3 * size_t(q);
      const float *T = N.q_T + 3 * size_t(q);
      c.x += T[0] * fmaxf(o[0], 0.f);
      c.y += T[1] * fmaxf(o[1], 0.f);
      c.z += T[2] * fmaxf(o[2], 0.f);
    }
    write_sample(F, i % F.width, i / F.width, accumID, c, 1e31f, true);
  }
}

// ---------------------------------------------------------------------------
// Field preparation
// ---------------------------------------------------------------------------
__global__ void k_bricks(const float *__restrict__ lin, int nx, int ny, int nz, int nbx, int nby,
                         int nbz, float *__restrict__ bricks) {
  // one thread per stored value: coalesced writes
  const size_t total = size_t(nbx) * nby * nbz * VRT_BRICK_FLOATS;
  for (size_t o = size_t(blockIdx.x) * blockDim.x + threadIdx.x; o < total;
       o += size_t(gridDim.x) * blockDim.x) {
    const size_t b = o / VRT_BRICK_FLOATS;
    const int l = int(o - b * VRT_BRICK_FLOATS);
    if (l >= 729) {  // record padding
      bricks[o] = 0.f;
      continue;
    }
    const int lx = l % 9, ly = (l / 9) % 9, lz = l / 81;
    const int bx = int(b % nbx), by = int((b / nbx) % nby), bz = int(b / (size_t(nbx) * nby));
    // the 9th voxel per axis overlaps the next brick; past the last voxel
    // the edge repeats (clamp addressing)
    const int x = min(bx * 8 + lx, nx - 1), y = min(by * 8 + ly, ny - 1),
              z = min(bz * 8 + lz, nz - 1);
    bricks[o] = lin[(size_t(z) * ny + y) * nx + x];
  }
}

// Value range of each macrocell: min/max over every voxel a trilinear sample
// inside the cell can read (see vrt_build_value_ranges).  One workgroup per
// macrocell.
__global__ void __launch_bounds__(256)
    k_value_ranges(KVolume v, float2 *__restrict__ ranges) {
  const int G0 = v.gdims[0], G1 = v.gdims[1], G2 = v.gdims[2];
  const int mc = blockIdx.x;
  const int m[3] = {mc % G0, (mc / G0) % G1, mc / (G0 * G1)};
  const int G[3] = {G0, G1, G2};
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    const double c0 = double(m[a]) * v.dims[a] / G[a] - 0.5;
    const double c1 = double(m[a] + 1) * v.dims[a] / G[a] - 0.5;
    lo[a] = max(0, int(floor(c0)));
    hi[a] = min(v.dims[a] - 1, int(floor(c1)) + 1);
  }
  const int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
  const int n = ex * ey * ez;
  float vmin = FLT_MAX, vmax = -FLT_MAX;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int x = lo[0] + i % ex, y = lo[1] + (i / ex) % ey, z = lo[2] + i / (ex * ey);
    const size_t brick = (size_t(z >> 3) * v.nby + size_t(y >> 3)) * v.nbx + size_t(x >> 3);
    const float val =
        v.bricks[brick * VRT_BRICK_FLOATS + size_t(((z & 7) * 9 + (y & 7)) * 9 + (x & 7))];
    vmin = fminf(vmin, val);
    vmax = fmaxf(vmax, val);
  }
  __shared__ float smin[256], smax[256];
  smin[threadIdx.x] = vmin;
  smax[threadIdx.x] = vmax;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) ranges[mc] = make_float2(smin[0], smax[0]);
}

// Atomic min / max of floats through their order-preserving integer images:
// non-negative floats order like their int bits, negative ones inversely to
// their unsigned bits.  -0 is folded into +0 first; NaN is ignored.
__device__ inline void atomic_min_float(float *a, float v) {
  if (v != v) return;
  v += 0.f;
  if (v >= 0.f)
    atomicMin(reinterpret_cast<int *>(a), __float_as_int(v));
  else
    atomicMax(reinterpret_cast<unsigned int *>(a), __float_as_uint(v));
}
__device__ inline void atomic_max_float(float *a, float v) {
  if (v != v) return;
  v += 0.f;
  if (v >= 0.f)
    atomicMax(reinterpret_cast<int *>(a), __float_as_int(v));
  else
    atomicMin(reinterpret_cast<unsigned int *>(a), __float_as_uint(v));
}

__global__ void __launch_bounds__(256) k_empty_ranges(float2 *__restrict__ r, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = make_float2(FLT_MAX, -FLT_MAX);
}

// unstructured: rasterizeBox of every tetrahedron's bounding box with its
// value range (UnstructuredField.cpp:387-399, GridAccel.h:51-67), the cells
// from projectOnGrid (GridAccel-common.h:18-24) over the grid bounds.  The
// reference's updateMC is a plain read-modify-write (a race when run in
// parallel); atomics make the result independent of the schedule.
__device__ inline int project_on_grid(float p, float lo, float hi, int G) {
  const float c = fminf(fmaxf((p - lo) / (hi - lo) * float(G), -1.f), float(G));
  return min(max(int(c), 0), G - 1);
}

__global__ void __launch_bounds__(256) k_tet_value_ranges(KVolume v, float2 *__restrict__ ranges) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= v.ug_num_tets) return;
  const float4 lo = v.ug_tets[2 * size_t(t)], hi = v.ug_tets[2 * size_t(t) + 1];
  const int G0 = v.gdims[0], G1 = v.gdims[1], G2 = v.gdims[2];
  const int x0 = project_on_grid(lo.x, v.glo.x, v.ghi.x, G0);
  const int x1 = project_on_grid(hi.x, v.glo.x, v.ghi.x, G0);
  const int y0 = project_on_grid(lo.y, v.glo.y, v.ghi.y, G1);
  const int y1 = project_on_grid(hi.y, v.glo.y, v.ghi.y, G1);
  const int z0 = project_on_grid(lo.z, v.glo.z, v.ghi.z, G2);
  const int z1 = project_on_grid(hi.z, v.glo.z, v.ghi.z, G2);
  for (int z = z0; z <= z1; ++z)
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) {
        float2 *r = ranges + (size_t(z) * G1 + y) * G0 + x;
        atomic_min_float(&r->x, lo.w);
        atomic_max_float(&r->y, hi.w);
      }
}

// nanovdb: value range of every block slot (a leaf's 512 values, or its
// constant)
__global__ void __launch_bounds__(256)
    k_nvdb_slot_ranges(const int32_t *__restrict__ table, size_t nslots,
                       const float *__restrict__ leaves, const float *__restrict__ consts,
                       float2 *__restrict__ out) {
  const size_t s = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const int e = table[s];
  if (e < 0) {
    const float c = consts[-1 - e];
    out[s] = make_float2(c, c);
    return;
  }
  const float4 *l = reinterpret_cast<const float4 *>(leaves + size_t(e) * 512u);
  float mn = FLT_MAX, mx = -FLT_MAX;
  for (int i = 0; i < 128; ++i) {
    const float4 q = l[i];
    mn = fminf(mn, fminf(fminf(q.x, q.y), fminf(q.z, q.w)));
    mx = fmaxf(mx, fmaxf(fmaxf(q.x, q.y), fmaxf(q.z, q.w)));
  }
  out[s] = make_float2(mn, mx);
}

// nanovdb: a macrocell's value range = union of the ranges of the 8^3 slots
// its voxel box touches (constant slots: their constant)
__global__ void __launch_bounds__(256)
    k_nvdb_value_ranges(KVolume v, int ncells, const int32_t *__restrict__ cell_voxels,
                        float2 *__restrict__ ranges) {
  const int c = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= ncells) return;
  const int32_t *b = cell_voxels + 6 * size_t(c);
  const float2 *slot_ranges = reinterpret_cast<const float2 *>(v.nv_slot_ranges);
  float mn = FLT_MAX, mx = -FLT_MAX;
  for (int sz = b[2] >> 3; sz <= (b[5] >> 3); ++sz)
    for (int sy = b[1] >> 3; sy <= (b[4] >> 3); ++sy)
      for (int sx = b[0] >> 3; sx <= (b[3] >> 3); ++sx) {
        const int tx = (sx >> 4) - v.nv_lo[0], ty = (sy >> 4) - v.nv_lo[1],
                  tz = (sz >> 4) - v.nv_lo[2];
        float lo, hi;
        int t = -1;
        if (uint32_t(tx) < uint32_t(v.nv_n[0]) && uint32_t(ty) < uint32_t(v.nv_n[1]) &&
            uint32_t(tz) < uint32_t(v.nv_n[2]))
          t = v.nv_top[(size_t(tz) * v.nv_n[1] + ty) * v.nv_n[0] + tx];
        if (t < 0) {
          lo = hi = v.nv_consts[-1 - t];
        } else {
          const float2 r =
              slot_ranges[size_t(t) * 4096u + uint32_t((((sz & 15) << 4) | (sy & 15)) << 4 | (sx & 15))];
          lo = r.x;
          hi = r.y;
        }
        mn = fminf(mn, lo);
        mx = fmaxf(mx, hi);
      }
  ranges[c] = make_float2(mn, mx);
}

// GridAccel::computeMaxOpacities (GridAccel.cpp:69-107) with the TF taps
// that tex1D can actually read for values in [vmin, vmax]
// The stored majorant is rounded up to a multiple q*step, q in 0..255, with
// step = (max TF opacity)/255 rounded up, so the grid fits in LDS as bytes
// (a larger majorant leaves delta tracking unbiased).
__global__ void __launch_bounds__(256)
    k_majorants(KVolume v, const float2 *__restrict__ ranges, int nmc, float *__restrict__ maj,
                uint8_t *__restrict__ bytes, float *__restrict__ step_out) {
  __shared__ float smax[256];
  float t = 0.f;
  for (int k = threadIdx.x; k < v.tf_size; k += blockDim.x) t = fmaxf(t, v.tf[k].w);
  smax[threadIdx.x] = t;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
    __syncthreads();
  }
  const float tf_max = smax[0];
  float step = tf_max / 255.f;
  if (step * 255.f < tf_max) step = __uint_as_float(__float_as_uint(step) + 1u);
  if (blockIdx.x == 0 && threadIdx.x == 0 && step_out) *step_out = step;

  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmc) return;
  const float2 r = ranges[i];
  float m = 0.f;
  if (!(r.y < r.x)) {
    const int n = v.tf_size;
    const float u0 = (r.x - v.vlo) / (v.vhi - v.vlo);
    const float u1 = (r.y - v.vlo) / (v.vhi - v.vlo);
    const float c0 = fminf(fmaxf(u0 * float(n) - 0.5f, -1.f), float(n));
    const float c1 = fminf(fmaxf(u1 * float(n) - 0.5f, -1.f), float(n));
    const int lo = min(max(int(floorf(c0)), 0), n - 1);
    const int hi = min(max(int(floorf(c1)) + 1, 0), n - 1);
    for (int k = lo; k <= hi; ++k) m = fmaxf(m, v.tf[k].w);
  }
  int q = 0;
  if (m > 0.f) {
    q = int(fminf(255.f, ceilf(m / step)));
    while (q < 255 && float(q) * step < m) ++q;
  }
  maj[i] = float(q) * step;
  if (bytes) bytes[i] = uint8_t(q);
}

__global__ void k_sample_field(KVolume v, int n, const float *__restrict__ pts,
                               float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float val;
  out[i] = sample_field(v, mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), val)
               ? val
               : __builtin_nanf("");
}

__global__ void __launch_bounds__(256)
    k_delta_track(KScene sc, int n, const float *__restrict__ rays, uint32_t *__restrict__ rng,
                  float *__restrict__ tout) {
  extern __shared__ float4 s_tf[];
  stage_tf(sc, s_tf);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *r = rays + 8 * i;
  KRay ray;
  ray.ori = mk(r[0], r[1], r[2]);
  ray.dir = mk(r[3], r[4], r[5]);
  ray.tmin = r[6];
  ray.tmax = r[7];
  uint32_t s = rng[i];
  uint32_t lookups = 0;
  VolHit h = delta_track(sc.vol[0], s_tf, lds_majorants(sc, s_tf, 0), lds_steps(sc, s_tf)[0],
                         to_local(sc.vol[0], ray), s, lookups);
  rng[i] = s;
  tout[i] = h.hit ? h.t : __builtin_inff();
}

__global__ void __launch_bounds__(256)
    k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16;
       i += size_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// closest / any surface hit of n rays {ori, dir, tmin, tmax}: t (inf on a
// miss) and the triangle's index in its mesh (-1)
__global__ void __launch_bounds__(256)
    k_intersect_surfaces(KScene sc, int n, const float *__restrict__ rays, int any,
                         float *__restrict__ t_out, int *__restrict__ prim_out,
                         int *__restrict__ surf_out) {
  extern __shared__ float4 s_tf[];  // [tables, unused here][stacks][staged nodes]
  stage_nodes(sc, s_tf);
  __syncthreads();
  int2 *s_bstack = reinterpret_cast<int2 *>(reinterpret_cast<char *>(s_tf) + scene_lds_bytes(sc));
  const float4 *lnodes = staged_nodes(sc, s_tf);
  const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const float *r = rays + 8 * size_t(i);
  KRay ray;
  ray.ori = mk(r[0], r[1], r[2]);
  ray.dir = mk(r[3], r[4], r[5]);
  ray.tmin = r[6];
  ray.tmax = r[7];
  bool hit = false;
  float best = ray.tmax;
  int prim = -1, surf = -1;
  for (int s = 0; s < sc.num_surfaces; ++s) {
    float t = best;
    int p = -1;
    uint32_t rec;
    if (bvh_hit(sc.surf[s], ray, any != 0, t, p, rec, s_bstack + threadIdx.x,
                lnodes + sc.staged[s].lds_first, sc.staged[s].lds_nodes)) {
      hit = true;
      best = t;
      prim = p;
      surf = s;
    }
  }
  if (sc.tls.num_inst > 0 && !(any && hit)) {
    float t = best;
    int slot = -1, p = -1;
    uint32_t rec;
    if (tls_hit(sc.tls, ray, any != 0, hit, t, slot, p, rec, s_bstack + threadIdx.x)) {
      hit = true;
      best = t;
      prim = p;
      surf = sc.num_surfaces + int(sc.tls.inst[slot].index);
    }
  }
  t_out[i] = hit ? best : __builtin_inff();
  prim_out[i] = hit ? prim : -1;
  if (surf_out) surf_out[i] = hit ? surf : -1;
}

__global__ void k_frame_reset(KFrame F) {
  const size_t n = size_t(F.width) * F.height;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    if (F.accum) F.accum[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (F.depth) F.depth[i] = 1e31f;
    if (F.accum64) {
      F.accum64[4 * i] = 0.0;
      F.accum64[4 * i + 1] = 0.0;
      F.accum64[4 * i + 2] = 0.0;
      F.accum64[4 * i + 3] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------


static bool scene_has_other_fields(const KScene &sc);

hipError_t launch_raycast(const KScene &sc, const KCamera &cam, const KParams &P, const KFrame &F,
                          uint32_t *queues, uint32_t *queues_next, int num_cus, hipStream_t s) {
  const bool surf = sc.num_surfaces > 0 || sc.tls.num_inst > 0;
  const size_t lds0 = scene_lds_bytes(sc) + (surf ? bvh_lds_stack_bytes() : 0);
  const bool any = scene_has_other_fields(sc), shade = P.gradient_shading != 0;
  using RaycastKernel = void (*)(KScene, KCamera, KParams, KFrame, uint32_t *, uint32_t *);
  static const RaycastKernel table[12] = {
      k_raycast<false, false, false, false>, k_raycast<true, false, false, false>,
      k_raycast<false, true, false, false>,  k_raycast<true, true, false, false>,
      k_raycast<false, false, true, false>,  k_raycast<true, false, true, false>,
      k_raycast<false, true, true, false>,   k_raycast<true, true, true, false>,
      k_raycast<false, false, true, true>,   k_raycast<true, false, true, true>,
      k_raycast<false, true, true, true>,    k_raycast<true, true, true, true>};
  // instanced meshes take the variants with the instance walk compiled in
  const bool inst = sc.tls.num_inst > 0;
  const RaycastKernel k =
      table[(inst ? 8 : surf ? 4 : 0) + (shade ? 2 : 0) + (any ? 1 : 0)];
  // one resident wave per SIMD slot, no more workgroups than tiles need
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, lds0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  KScene scs = sc;
  if (surf) plan_staged_nodes(scs, lds0, per_cu);
  const size_t lds = lds0 + staged_nodes_bytes(scs);
  const int ntiles = ((F.width + 7) / 8) * ((F.height + 7) / 8);
  int blocks = std::min(std::max(1, num_cus) * per_cu, (ntiles + 3) / 4);
  blocks = std::max(8, (blocks + 7) / 8 * 8);  // every XCD's queue gets workgroups
  hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, s, scs, cam, P, F, queues, queues_next);
  return hipGetLastError();
}

// scenes with surfaces or quad lights take the kernel with the BVH code
static bool scene_has_area(const KScene &sc) {
  if (sc.num_surfaces > 0 || sc.tls.num_inst > 0) return true;
  for (int i = 0; i < sc.num_lights; ++i)
    if (sc.light[i].type == VRT_LIGHT_QUAD || sc.light[i].type == VRT_LIGHT_HDRI) return true;
  return false;
}

// scenes with a nanovdb or unstructured field take the kernel with the
// field-type dispatch
static bool scene_has_other_fields(const KScene &sc) {
  for (int i = 0; i < sc.num_volumes; ++i)
    if (sc.vol[i].field_type != VRT_FIELD_STRUCTURED_REGULAR) return true;
  return false;
}

using PathtraceKernel = void (*)(KScene, KCamera, KParams, KFrame, int, uint32_t *, uint32_t *,
                                 KNrc);

static PathtraceKernel pathtrace_kernel(const KScene &sc, bool nrc, bool grad) {
  static const PathtraceKernel table[16] = {
      k_pathtrace<false, false, false, false, false>, k_pathtrace<false, false, true, false, false>,
      k_pathtrace<false, true, false, false, false>,  k_pathtrace<false, true, true, false, false>,
      k_pathtrace<true, false, false, false, false>,  k_pathtrace<true, false, true, false, false>,
      k_pathtrace<true, true, false, false, false>,   k_pathtrace<true, true, true, false, false>,
      k_pathtrace<false, false, false, true, false>,  k_pathtrace<false, false, true, true, false>,
      k_pathtrace<false, true, false, true, false>,   k_pathtrace<false, true, true, true, false>,
      k_pathtrace<true, false, false, true, false>,   k_pathtrace<true, false, true, true, false>,
      k_pathtrace<true, true, false, true, false>,    k_pathtrace<true, true, true, true, false>};
  // instanced meshes: the surface variants with the instance walk compiled in
  static const PathtraceKernel inst_table[8] = {
      k_pathtrace<false, true, false, false, true>, k_pathtrace<false, true, true, false, true>,
      k_pathtrace<true, true, false, false, true>,  k_pathtrace<true, true, true, false, true>,
      k_pathtrace<false, true, false, true, true>,  k_pathtrace<false, true, true, true, true>,
      k_pathtrace<true, true, false, true, true>,   k_pathtrace<true, true, true, true, true>};
  if (sc.tls.num_inst > 0)
    return inst_table[(grad ? 4 : 0) + (nrc ? 2 : 0) + (scene_has_other_fields(sc) ? 1 : 0)];
  return table[(grad ? 8 : 0) + (nrc ? 4 : 0) + (scene_has_area(sc) ? 2 : 0) +
               (scene_has_other_fields(sc) ? 1 : 0)];
}

// Pass scheduling of the path tracer (vrt_params.pass_schedule): 1 chains the
// passes of a bounce (KParams::chain_passes), 2 traces one ray per lane and
// pass, 3 is 2 with idle lanes refilled only 16 at a time (refill_min).
// Measured frame times (1 / 2 / 3): fBm 512^3 1.21 / 1.27 / 1.26 ms, mixed
// surfaces + volume 18.6 / 19.5 / 19.4 ms, NanoVDB cloud 1024^3 3.44 / 3.59 /
// 3.36 ms (its cache frames 2.82 / 2.83 / 2.77 ms), 2.06 M tetrahedra at
// 3840x2160 22.0 / 20.5 / 20.9 ms; the default picks per scene: 3 with a
// NanoVDB field, 2 with a tetrahedral one (and no NanoVDB field), 1
// otherwise.
int pathtrace_schedule(const KScene &sc) {
  bool nvdb = false, tets = false;
  for (int i = 0; i < sc.num_volumes; ++i) {
    nvdb = nvdb || sc.vol[i].field_type == VRT_FIELD_NANOVDB;
    tets = tets || sc.vol[i].field_type == VRT_FIELD_UNSTRUCTURED;
  }
  return nvdb ? 3 : tets ? 2 : 1;
}

void set_pass_schedule(KParams &P, int schedule) {
  P.chain_passes = schedule == 1;
  P.refill_min = schedule == 3 ? 16 : 1;
}

size_t pathtrace_lds_size(const KScene &sc) {
  return scene_lds_bytes(sc) +
         (scene_has_area(sc) ? bvh_lds_stack_bytes() + staged_nodes_bytes(sc) : 0);
}

// The top BVH nodes each workgroup stages: the first nodes of surface 0's
// breadth-first array, then surface 1's, ..., as many as fit beside `base`
// LDS bytes per workgroup without lowering the kernel's `per_cu` resident
// workgroups (1 KB kept for allocation granularity), at most kMaxStagedNodes.
void plan_staged_nodes(KScene &sc, size_t base, int per_cu) {
  for (int i = 0; i < sc.num_surfaces; ++i) sc.staged[i].lds_nodes = sc.staged[i].lds_first = 0;
  const size_t per_wg = kLdsBytesPerCu / size_t(std::max(1, per_cu));
  if (per_wg <= base + 1024) return;
  size_t room = std::min((per_wg - base - 1024) / (8 * sizeof(float4)), size_t(kMaxStagedNodes));
  int first = 0;
  for (int i = 0; i < sc.num_surfaces && room > 0; ++i) {
    const int k = int(std::min(room, size_t(std::max(0, sc.staged[i].num_nodes))));
    sc.staged[i].lds_nodes = k;
    sc.staged[i].lds_first = first;
    first += 8 * k;
    room -= size_t(k);
  }
}

hipError_t launch_pathtrace(const KScene &sc, const KCamera &cam, const KParams &P,
                            const KFrame &F, int accumID, uint32_t *queues,
                            uint32_t *queues_next, int blocks, hipStream_t s, const KNrc *nrc) {
  const size_t lds = pathtrace_lds_size(sc);
  if (nrc) hipLaunchKernelGGL(k_nrc_frame_begin, dim3(1), dim3(64), 0, s, *nrc);
  hipLaunchKernelGGL(pathtrace_kernel(sc, nrc != nullptr, P.gradient_shading != 0), dim3(blocks),
                     dim3(256), lds, s, sc,
                     cam, P, F, accumID, queues, queues_next, nrc ? *nrc : KNrc{});
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Work-queue counters of vrt_pathtrace_noalloc: a file-scope device array
// (two sets of kQueues words, 64 B) in place of the context's allocation.
// Every GPU holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns on that device.
// Each call clears the set it claims from with a kernel (no memset node), so
// the call may be captured into a graph; calls that overlap on one GPU share
// the array and race.
// ---------------------------------------------------------------------------
constexpr int kScratchWords = 2 * kQueues;
static __device__ __align__(256) uint32_t scratch[kScratchWords];

__global__ void k_clear_words(uint32_t *p, int n) {
  if (int(threadIdx.x) < n) p[threadIdx.x] = 0u;
}

hipError_t static_queue_address(uint32_t **ptr) {
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) *ptr = static_cast<uint32_t *>(p);
  return e;
}

hipError_t launch_pathtrace_static(const KScene &sc, const KCamera &cam, const KParams &P,
                                   const KFrame &F, int accumID, uint32_t *scratch_ptr,
                                   int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, s, scratch_ptr, kScratchWords);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_pathtrace(sc, cam, P, F, accumID, scratch_ptr, scratch_ptr + kQueues, blocks, s);
}

hipError_t launch_taa_resolve(const KFrame &F, float alpha, hipStream_t s) {
  const int n = F.width * F.height;
  hipLaunchKernelGGL(k_taa_resolve, dim3((n + 255) / 256), dim3(256), 0, s, F, alpha);
  hipError_t e = hipGetLastError();
  // the history for the next frame (Frame.cpp:362-380)
  if (e == hipSuccess)
    e = hipMemcpyAsync(F.taa_prev, F.taa_curr, size_t(n) * sizeof(float4), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(F.taa_prev_alb, F.taa_curr_alb, size_t(n) * 12, hipMemcpyDeviceToDevice, s);
  return e;
}

hipError_t launch_nrc_train_targets(const KNrc &N, hipStream_t s) {
  if (N.t_cap <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nrc_train_targets, dim3(std::min((N.t_cap + 255) / 256, 4096)), dim3(256),
                     0, s, N);
  return hipGetLastError();
}

hipError_t launch_nrc_resolve(const KFrame &F, const KNrc &N, int accumID, hipStream_t s) {
  const int n = F.width * F.height;
  hipLaunchKernelGGL(k_nrc_resolve, dim3(std::min((n + 255) / 256, 65536)), dim3(256), 0, s, F,
                     N, accumID);
  return hipGetLastError();
}

int pathtrace_blocks_per_cu(const KScene &sc, size_t lds_bytes, bool nrc, bool grad) {
  int n = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &n, pathtrace_kernel(sc, nrc, grad), 256, lds_bytes);
  if (e != hipSuccess) return 1;
  return n > 0 ? n : 1;
}

hipError_t launch_bricks(const float *lin, const int dims[3], float *bricks, hipStream_t s) {
  const int nbx = (dims[0] + 7) / 8, nby = (dims[1] + 7) / 8, nbz = (dims[2] + 7) / 8;
  const size_t total = size_t(nbx) * nby * nbz * VRT_BRICK_FLOATS;
  const size_t blocks = std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_bricks, dim3(unsigned(blocks)), dim3(256), 0, s, lin, dims[0], dims[1],
                     dims[2], nbx, nby, nbz, bricks);
  return hipGetLastError();
}

hipError_t launch_value_ranges(const KVolume &v, float *ranges, hipStream_t s) {
  const int nmc = v.gdims[0] * v.gdims[1] * v.gdims[2];
  hipLaunchKernelGGL(k_value_ranges, dim3(nmc), dim3(256), 0, s, v,
                     reinterpret_cast<float2 *>(ranges));
  return hipGetLastError();
}

size_t scene_lds_size(const KScene &sc) { return scene_lds_bytes(sc); }

hipError_t launch_nanovdb_slot_ranges(const int32_t *table, size_t nslots, const float *leaves,
                                      const float *consts, float *slot_ranges, hipStream_t s) {
  if (nslots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_nvdb_slot_ranges, dim3(unsigned((nslots + 255) / 256)), dim3(256), 0, s,
                     table, nslots, leaves, consts, reinterpret_cast<float2 *>(slot_ranges));
  return hipGetLastError();
}

hipError_t launch_nanovdb_value_ranges(const KVolume &v, const int32_t *cell_voxels,
                                       float *ranges, hipStream_t s) {
  const int n = v.gdims[0] * v.gdims[1] * v.gdims[2];
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nvdb_value_ranges, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, v, n,
                     cell_voxels, reinterpret_cast<float2 *>(ranges));
  return hipGetLastError();
}

hipError_t launch_tet_value_ranges(const KVolume &v, float *ranges, hipStream_t s) {
  const int n = v.gdims[0] * v.gdims[1] * v.gdims[2];
  if (n <= 0) return hipSuccess;
  float2 *r = reinterpret_cast<float2 *>(ranges);
  hipLaunchKernelGGL(k_empty_ranges, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, r, n);
  if (v.ug_num_tets)
    hipLaunchKernelGGL(k_tet_value_ranges, dim3((v.ug_num_tets + 255u) / 256u), dim3(256), 0, s,
                       v, r);
  return hipGetLastError();
}

hipError_t launch_majorants(const KVolume &v, const float *ranges, float *maj, uint8_t *bytes,
                            float *step, hipStream_t s) {
  const int nmc = v.gdims[0] * v.gdims[1] * v.gdims[2];
  hipLaunchKernelGGL(k_majorants, dim3((nmc + 255) / 256), dim3(256), 0, s, v,
                     reinterpret_cast<const float2 *>(ranges), nmc, maj, bytes, step);
  return hipGetLastError();
}

hipError_t launch_sample_field(const KVolume &v, int n, const float *pts, float *out,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_sample_field, dim3((n + 255) / 256), dim3(256), 0, s, v, n, pts, out);
  return hipGetLastError();
}

hipError_t launch_intersect_surfaces(const KScene &sc, int n, const float *rays, int any,
                                     float *t_out, int *prim_out, int *surf_out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const size_t lds0 = staged_nodes_offset(sc);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_intersect_surfaces, kBvhLanes,
                                                   lds0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  KScene scs = sc;
  plan_staged_nodes(scs, lds0, per_cu);
  hipLaunchKernelGGL(k_intersect_surfaces, dim3((n + 255) / 256), dim3(kBvhLanes),
                     lds0 + staged_nodes_bytes(scs), s, scs, n, rays, any, t_out, prim_out,
                     surf_out);
  return hipGetLastError();
}

hipError_t launch_delta_track(const KScene &sc, int n, const float *rays, uint32_t *rng,
                              float *tout, hipStream_t s) {
  const size_t lds = scene_lds_bytes(sc);
  hipLaunchKernelGGL(k_delta_track, dim3((n + 255) / 256), dim3(256), lds, s, sc, n, rays, rng,
                     tout);
  return hipGetLastError();
}

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
  const size_t n16 = bytes / 16;
  const size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 64);
  if (n16)
    hipLaunchKernelGGL(k_copy, dim3(unsigned(blocks)), dim3(256), 0, s,
                       reinterpret_cast<const uint4 *>(src), reinterpret_cast<uint4 *>(dst), n16);
  return hipGetLastError();
}

hipError_t launch_frame_reset(const KFrame &F, hipStream_t s) {
  const size_t n = size_t(F.width) * F.height;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_frame_reset, dim3(unsigned(blocks)), dim3(256), 0, s, F);
  return hipGetLastError();
}

}  // namespace vrt
