// vrt_bvh_build_gpu: the triangle BVH built on the GPU, for meshes that
// change between frames (the reference builds BLS BVHs on the host and
// uploads them, DeviceBVH.h:153-255, 320; its CUDA path uses visionaray's
// LBVH builder for element BVHs, UnstructuredField.cpp:243-256).
//
// Linear BVH (Karras, "Maximizing Parallelism in the Construction of BVHs,
// Octrees, and k-d Trees", HPG 2012): 63-bit Morton codes of the triangle
// box centres over their bounds, a radix sort (hipCUB), one thread per inner
// node finding its key range and split from common-prefix lengths (ties
// between equal codes broken by index), boxes refitted bottom-up one launch
// per tree level (nodes bucketed by depth, deepest level first, so kernel
// boundaries order the box writes), then a top-down
// collapse into the 4-wide node layout of bvh_build.cpp, one launch per
// level: a node opens its largest child while it has fewer than 4, and any
// subtree of at most 4 triangles -- a contiguous run in Morton order --
// becomes a leaf.  Triangles are stored in Morton order.  Hits are the
// same as with the host-built BVH (closest hit, ties to the lower triangle
// index); only the tree, and so the traversal cost, differs.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cstring>
#include <vector>

#include "abi_internal.h"
#include "bvh_internal.h"
#include "vrt.h"

namespace {

using namespace vrt;

// float -> unsigned, monotone (for atomic min/max of floats)
__device__ inline unsigned ord_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float from_ord(unsigned u) {
  const unsigned b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

// triangles in input order as {v1, prim id bits}, e1, e2 (the form the hit
// kernels read), their boxes, box centres, and the centres' bounds
__global__ void __launch_bounds__(256)
    k_bvh_prep(const float *__restrict__ vtx, int64_t nv, const uint32_t *__restrict__ idx,
               int64_t n, float4 *__restrict__ tris, float *__restrict__ box,
               float *__restrict__ cen, unsigned *__restrict__ bounds, unsigned *err) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  float c[3] = {0.f, 0.f, 0.f};
  bool ok = i < n;
  if (ok) {
    const uint32_t a = idx ? idx[3 * i] : uint32_t(3 * i);
    const uint32_t b = idx ? idx[3 * i + 1] : uint32_t(3 * i + 1);
    const uint32_t d = idx ? idx[3 * i + 2] : uint32_t(3 * i + 2);
    if (int64_t(a) >= nv || int64_t(b) >= nv || int64_t(d) >= nv) {
      atomicOr(err, 1u);
      ok = false;
    } else {
      const float *p0 = vtx + 3 * size_t(a), *p1 = vtx + 3 * size_t(b), *p2 = vtx + 3 * size_t(d);
      float lo[3], hi[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(p0[k], fminf(p1[k], p2[k]));
        hi[k] = fmaxf(p0[k], fmaxf(p1[k], p2[k]));
        c[k] = 0.5f * (lo[k] + hi[k]);
        box[6 * i + k] = lo[k];
        box[6 * i + 3 + k] = hi[k];
        cen[3 * i + k] = c[k];
      }
      tris[3 * i] = make_float4(p0[0], p0[1], p0[2], __uint_as_float(uint32_t(i)));
      tris[3 * i + 1] = make_float4(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2], 0.f);
      tris[3 * i + 2] = make_float4(p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2], 0.f);
    }
  }
  // workgroup min/max of the centres, then one atomic per workgroup and bound
  __shared__ unsigned smin[3][256], smax[3][256];
  for (int k = 0; k < 3; ++k) {
    smin[k][threadIdx.x] = ok ? ord_bits(c[k]) : 0xFFFFFFFFu;
    smax[k][threadIdx.x] = ok ? ord_bits(c[k]) : 0u;
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s)
      for (int k = 0; k < 3; ++k) {
        smin[k][threadIdx.x] = min(smin[k][threadIdx.x], smin[k][threadIdx.x + s]);
        smax[k][threadIdx.x] = max(smax[k][threadIdx.x], smax[k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) {
      atomicMin(&bounds[k], smin[k][0]);
      atomicMax(&bounds[3 + k], smax[k][0]);
    }
}

// 21 bits of v spread to every third bit
__device__ inline uint64_t spread3(uint32_t v) {
  uint64_t x = 0;
  for (int b = 0; b < 21; ++b) x |= uint64_t((v >> b) & 1u) << (3 * b);
  return x;
}

__global__ void __launch_bounds__(256)
    k_bvh_morton(int64_t n, const float *__restrict__ cen, float3 lo, float3 ext,
                 uint64_t *__restrict__ keys, uint32_t *__restrict__ ids) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const float e[3] = {ext.x, ext.y, ext.z}, l[3] = {lo.x, lo.y, lo.z};
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const float u = e[k] > 0.f ? (cen[3 * i + k] - l[k]) / e[k] : 0.f;
    q[k] = uint32_t(fminf(fmaxf(u, 0.f), 1.f) * 2097151.f);
  }
  keys[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
  ids[i] = uint32_t(i);
}

// length of the common prefix of keys i and j (-1 outside [0, n)); equal
// keys compare their indices
__device__ inline int prefix(const uint64_t *__restrict__ k, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = k[i], b = k[j];
  if (a == b) return 64 + __clz(int(uint32_t(i ^ j)));
  return __clzll((long long)(a ^ b));
}

// inner node i: children (>= 0 inner node, < 0: leaf -1-k), key range
__global__ void __launch_bounds__(256)
    k_bvh_hierarchy(int64_t n, const uint64_t *__restrict__ keys, int *__restrict__ left,
                    int *__restrict__ right, int *__restrict__ first, int *__restrict__ last,
                    int *__restrict__ parent, int *__restrict__ leaf_parent) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n - 1) return;
  const int pr = prefix(keys, n, i, i + 1), pl = prefix(keys, n, i, i - 1);
  const int64_t d = pr > pl ? 1 : -1;
  const int pmin = prefix(keys, n, i, i - d);
  int64_t lmax = 2;
  while (prefix(keys, n, i, i + lmax * d) > pmin) lmax <<= 1;
  int64_t l = 0;
  for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
    if (prefix(keys, n, i, i + (l + t) * d) > pmin) l += t;
  const int64_t j = i + l * d;
  const int pnode = prefix(keys, n, i, j);
  int64_t s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (prefix(keys, n, i, i + (s + t) * d) > pnode) s += t;
  } while (t > 1);
  const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
  const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
  const int L = lo == gamma ? int(-1 - gamma) : int(gamma);
  const int R = hi == gamma + 1 ? int(-2 - gamma) : int(gamma + 1);
  left[i] = L;
  right[i] = R;
  first[i] = int(lo);
  last[i] = int(hi);
  if (L >= 0) parent[L] = int(i); else leaf_parent[-1 - L] = int(i);
  if (R >= 0) parent[R] = int(i); else leaf_parent[-1 - R] = int(i);
}

__device__ inline void child_box(int c, const uint32_t *__restrict__ ids,
                                 const float *__restrict__ tbox, const float *nbox, float b[6]) {
  if (c < 0) {
    const float *t = tbox + 6 * size_t(ids[-1 - c]);
    for (int k = 0; k < 6; ++k) b[k] = t[k];
  } else {
    for (int k = 0; k < 6; ++k) b[k] = nbox[6 * size_t(c) + k];
  }
}

// A child's key range shares a longer common prefix than its parent's (at
// most 63 key bits plus 32 index bits), so the binary tree is at most 97
// levels deep.
constexpr int kMaxLevels = 128;

// depth of each inner node (root 0) by walking its parent chain, and a
// histogram of the depths (per workgroup in LDS, then one atomic per level)
__global__ void __launch_bounds__(256)
    k_bvh_depth(int64_t ni, const int *__restrict__ parent, uint8_t *__restrict__ depth,
                unsigned *__restrict__ level_count, unsigned *__restrict__ err) {
  __shared__ unsigned hist[kMaxLevels];
  for (int b = threadIdx.x; b < kMaxLevels; b += 256) hist[b] = 0;
  __syncthreads();
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < ni) {
    int d = 0;
    for (int p = parent[i]; p >= 0 && d < kMaxLevels; p = parent[p]) ++d;
    if (d >= kMaxLevels) {
      atomicOr(err, 1u);
      d = kMaxLevels - 1;
    }
    depth[i] = uint8_t(d);
    atomicAdd(&hist[d], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kMaxLevels; b += 256)
    if (hist[b]) atomicAdd(&level_count[b], hist[b]);
}

// inner nodes listed level by level: cursor[d] starts at level d's offset;
// a workgroup reserves its run of each level with one atomic
__global__ void __launch_bounds__(256)
    k_bvh_bucket(int64_t ni, const uint8_t *__restrict__ depth, unsigned *__restrict__ cursor,
                 int *__restrict__ order) {
  __shared__ unsigned hist[kMaxLevels], base[kMaxLevels];
  for (int b = threadIdx.x; b < kMaxLevels; b += 256) hist[b] = 0;
  __syncthreads();
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  unsigned d = 0, local = 0;
  if (i < ni) {
    d = depth[i];
    local = atomicAdd(&hist[d], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kMaxLevels; b += 256)
    if (hist[b]) base[b] = atomicAdd(&cursor[b], hist[b]);
  __syncthreads();
  if (i < ni) order[base[d] + local] = int(i);
}

// boxes of one level's inner nodes from their children's (the deeper
// levels were written by earlier launches)
__global__ void __launch_bounds__(256)
    k_bvh_refit_level(int count, const int *__restrict__ nodes, const uint32_t *__restrict__ ids,
                      const float *__restrict__ tbox, const int *__restrict__ left,
                      const int *__restrict__ right, float *nbox) {
  const int w = blockIdx.x * 256 + threadIdx.x;
  if (w >= count) return;
  const int p = nodes[w];
  float a[6], b[6];
  child_box(left[p], ids, tbox, nbox, a);
  child_box(right[p], ids, tbox, nbox, b);
  for (int q = 0; q < 3; ++q) {
    nbox[6 * size_t(p) + q] = fminf(a[q], b[q]);
    nbox[6 * size_t(p) + 3 + q] = fmaxf(a[3 + q], b[3 + q]);
  }
}

// one level of the 4-wide collapse: work item = (inner node, 4-wide slot)
__global__ void __launch_bounds__(256)
    k_bvh_collapse(int count, const int2 *__restrict__ work, const int *__restrict__ left,
                   const int *__restrict__ right, const int *__restrict__ first,
                   const int *__restrict__ last, const uint32_t *__restrict__ ids,
                   const float *__restrict__ tbox, const float *__restrict__ nbox,
                   float4 *__restrict__ nodes, int2 *__restrict__ next,
                   unsigned *__restrict__ next_count, unsigned *__restrict__ node_count) {
  const int w = blockIdx.x * 256 + threadIdx.x;
  if (w >= count) return;
  const int root = work[w].x, slot = work[w].y;
  auto size = [&](int c) { return c < 0 ? 1 : last[c] - first[c] + 1; };
  auto area = [&](int c) {
    float b[6];
    child_box(c, ids, tbox, nbox, b);
    const float x = b[3] - b[0], y = b[4] - b[1], z = b[5] - b[2];
    return 2.f * (x * y + y * z + z * x);
  };
  int kids[4];
  int nk = 0;
  if (root < 0) {
    kids[nk++] = root;  // a single triangle
  } else {
    kids[nk++] = left[root];
    kids[nk++] = right[root];
    while (nk < 4) {
      int best = -1;
      float ba = -1.f;
      for (int k = 0; k < nk; ++k)
        if (kids[k] >= 0 && size(kids[k]) > 4) {
          const float a = area(kids[k]);
          if (a > ba) {
            ba = a;
            best = k;
          }
        }
      if (best < 0) break;
      const int c = kids[best];
      kids[best] = left[c];
      kids[nk++] = right[c];
    }
  }
  float lanes[8][4];
  for (int k = 0; k < 4; ++k) {
    lanes[0][k] = lanes[2][k] = lanes[4][k] = 1.f;  // empty slot: inverted box
    lanes[1][k] = lanes[3][k] = lanes[5][k] = -1.f;
    int e = -1;
    unsigned cnt = 0;
    if (k < nk) {
      const int c = kids[k];
      float b[6];
      child_box(c, ids, tbox, nbox, b);
      for (int q = 0; q < 3; ++q) {
        lanes[2 * q][k] = b[q];
        lanes[2 * q + 1][k] = b[3 + q];
      }
      if (c < 0) {
        e = -1 - c;
        cnt = 1;
      } else if (size(c) <= 4) {
        e = first[c];
        cnt = unsigned(size(c));
      } else {
        e = int(atomicAdd(node_count, 1u));
        next[atomicAdd(next_count, 1u)] = make_int2(c, e);
      }
    }
    lanes[6][k] = __int_as_float(e);
    lanes[7][k] = __uint_as_float(cnt);
  }
  for (int r = 0; r < 8; ++r)
    nodes[8 * size_t(slot) + r] = make_float4(lanes[r][0], lanes[r][1], lanes[r][2], lanes[r][3]);
}

__global__ void __launch_bounds__(256)
    k_bvh_order_tris(int64_t n, const uint32_t *__restrict__ ids, const float4 *__restrict__ in,
                     float4 *__restrict__ out) {
  const int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (k >= n) return;
  const size_t t = ids[k];
  out[3 * k] = in[3 * t];
  out[3 * k + 1] = in[3 * t + 1];
  out[3 * k + 2] = in[3 * t + 2];
}

// the build's scratch buffers, carved from one device allocation (one
// hipMalloc and one hipFree per build instead of one per buffer)
struct Scratch {
  void *base = nullptr;
  std::vector<std::pair<void **, size_t>> parts;
  template <class T>
  void add(T **p, size_t bytes) {
    parts.emplace_back(reinterpret_cast<void **>(p), (std::max<size_t>(bytes, 16) + 255) & ~size_t(255));
  }
  hipError_t commit() {
    size_t total = 0;
    for (auto &q : parts) total += q.second;
    hipError_t e = hipMalloc(&base, total);
    if (e != hipSuccess) return e;
    char *c = static_cast<char *>(base);
    for (auto &q : parts) {
      *q.first = c;
      c += q.second;
    }
    return hipSuccess;
  }
  ~Scratch() {
    if (base) (void)hipFree(base);
  }
};

}  // namespace

extern "C" {

int vrt_bvh_build_gpu(vrt_context ctx, const float *vertices, int64_t num_vertices,
                      const uint32_t *indices, int64_t num_triangles, vrt_bvh *out) {
  if (int r = check_ctx(ctx)) return r;
  if (!out || !vertices || num_triangles < 1 || num_vertices < 3)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad bvh arguments");
  if (num_triangles > (int64_t(1) << 30)) return fail(VRT_ERROR_UNSUPPORTED, "too many triangles");
  if (!indices && num_vertices < 3 * num_triangles)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "triangle soup needs 3 vertices per triangle");
  *out = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  const int64_t n = num_triangles;
  const unsigned blocks = unsigned((n + 255) / 256);
  Scratch S;
  float4 *tris_in = nullptr;
  float *tbox = nullptr, *cen = nullptr, *nbox = nullptr;
  unsigned *bounds = nullptr, *err = nullptr, *levels = nullptr, *counters = nullptr;
  uint8_t *depth8 = nullptr;
  int *order = nullptr;
  uint64_t *keys = nullptr, *keys_sorted = nullptr;
  uint32_t *ids = nullptr, *ids_sorted = nullptr;
  int *left = nullptr, *right = nullptr, *first = nullptr, *last = nullptr, *parent = nullptr,
      *leaf_parent = nullptr;
  int2 *work[2] = {nullptr, nullptr};
  const int64_t ni = std::max<int64_t>(n - 1, 1);
  size_t temp_bytes = 0;
  void *temp = nullptr;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys_sorted, ids,
                                                    ids_sorted, int(n), 0, 63, s);
  if (e != hipSuccess) return hip_fail(e, "bvh sort");
  S.add(&tris_in, size_t(n) * 48);
  S.add(&tbox, size_t(n) * 24);
  S.add(&cen, size_t(n) * 12);
  S.add(&bounds, 6 * 4);
  S.add(&err, 4);
  S.add(&keys, size_t(n) * 8);
  S.add(&keys_sorted, size_t(n) * 8);
  S.add(&ids, size_t(n) * 4);
  S.add(&ids_sorted, size_t(n) * 4);
  S.add(&temp, temp_bytes);
  S.add(&left, size_t(ni) * 4);
  S.add(&right, size_t(ni) * 4);
  S.add(&first, size_t(ni) * 4);
  S.add(&last, size_t(ni) * 4);
  S.add(&parent, size_t(ni) * 4);
  S.add(&leaf_parent, size_t(n) * 4);
  S.add(&nbox, size_t(ni) * 24);
  S.add(&levels, 2 * kMaxLevels * 4);
  S.add(&depth8, size_t(ni));
  S.add(&order, size_t(ni) * 4);
  S.add(&counters, 8);
  S.add(&work[0], size_t(n) * 8);
  S.add(&work[1], size_t(n) * 8);
  e = S.commit();
  if (e != hipSuccess) return hip_fail(e, "bvh scratch");
  const unsigned init_bounds[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
  e = hipMemcpyAsync(bounds, init_bounds, sizeof(init_bounds), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(err, 0, 4, s);
  if (e != hipSuccess) return hip_fail(e, "bvh init");
  hipLaunchKernelGGL(k_bvh_prep, dim3(blocks), dim3(256), 0, s, vertices, num_vertices, indices,
                     n, tris_in, tbox, cen, bounds, err);
  unsigned hb[7];
  e = hipMemcpyAsync(hb, bounds, 24, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(hb + 6, err, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "bvh prepare");
  if (hb[6]) return fail(VRT_ERROR_INVALID_ARGUMENT, "triangle index out of range");
  float3 lo = make_float3(from_ord(hb[0]), from_ord(hb[1]), from_ord(hb[2]));
  float3 ext = make_float3(from_ord(hb[3]) - lo.x, from_ord(hb[4]) - lo.y, from_ord(hb[5]) - lo.z);
  hipLaunchKernelGGL(k_bvh_morton, dim3(blocks), dim3(256), 0, s, n, cen, lo, ext, keys, ids);
  e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_sorted, ids, ids_sorted,
                                         int(n), 0, 63, s);
  if (e != hipSuccess) return hip_fail(e, "bvh sort");

  // parent of the root and of a lone leaf: -1
  e = hipMemsetAsync(parent, 0xFF, size_t(ni) * 4, s);
  if (e == hipSuccess) e = hipMemsetAsync(leaf_parent, 0xFF, size_t(n) * 4, s);
  if (e == hipSuccess) e = hipMemsetAsync(levels, 0, kMaxLevels * 4, s);
  if (e == hipSuccess) e = hipMemsetAsync(err, 0, 4, s);
  if (e != hipSuccess) return hip_fail(e, "bvh init");
  if (n > 1) {
    const unsigned iblocks = unsigned((ni + 255) / 256);
    hipLaunchKernelGGL(k_bvh_hierarchy, dim3(iblocks), dim3(256), 0, s, n, keys_sorted, left,
                       right, first, last, parent, leaf_parent);
    // bottom-up boxes, one launch per level from the deepest to the root
    hipLaunchKernelGGL(k_bvh_depth, dim3(iblocks), dim3(256), 0, s, ni, parent, depth8, levels,
                       err);
    unsigned hl[kMaxLevels + 1];
    e = hipMemcpyAsync(hl, levels, kMaxLevels * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hl + kMaxLevels, err, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "bvh levels");
    if (hl[kMaxLevels]) return fail(VRT_ERROR_UNSUPPORTED, "bvh deeper than its key bits allow");
    unsigned off[kMaxLevels];
    unsigned sum = 0;
    for (int d = 0; d < kMaxLevels; ++d) {
      off[d] = sum;
      sum += hl[d];
    }
    if (sum != unsigned(ni)) return fail(VRT_ERROR_UNSUPPORTED, "bvh level counts do not add up");
    e = hipMemcpyAsync(levels + kMaxLevels, off, sizeof(off), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "bvh levels");
    hipLaunchKernelGGL(k_bvh_bucket, dim3(iblocks), dim3(256), 0, s, ni, depth8,
                       levels + kMaxLevels, order);
    for (int d = kMaxLevels - 1; d >= 0; --d)
      if (hl[d])
        hipLaunchKernelGGL(k_bvh_refit_level, dim3((hl[d] + 255) / 256), dim3(256), 0, s,
                           int(hl[d]), order + off[d], ids_sorted, tbox, left, right, nbox);
    // the host copy of the offsets must outlive the asynchronous upload
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "bvh refit");
  }

  auto *h = new vrt_bvh_s;
  h->device = ctx->device;
  h->num_tris = uint32_t(n);
  e = hipMalloc(&h->nodes, size_t(ni + 1) * 8 * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&h->tris, size_t(n) * 3 * sizeof(float4));
  auto cleanup = [&](hipError_t err_code, const char *what) {
    (void)hipFree(h->nodes);
    (void)hipFree(h->tris);
    delete h;
    return hip_fail(err_code, what);
  };
  if (e != hipSuccess) return cleanup(e, "bvh allocation");
  hipLaunchKernelGGL(k_bvh_order_tris, dim3(blocks), dim3(256), 0, s, n, ids_sorted, tris_in,
                     h->tris);
  // collapse, level by level from the root (4-wide node 0)
  const int2 root = make_int2(n > 1 ? 0 : -1, 0);
  const unsigned one[2] = {0u, 1u};  // [0] next-level count, [1] nodes allocated
  e = hipMemcpyAsync(work[0], &root, sizeof(root), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(counters, one, 8, hipMemcpyHostToDevice, s);
  unsigned count = 1, depth = 0, cur = 0;
  while (e == hipSuccess && count > 0) {
    e = hipMemsetAsync(counters, 0, 4, s);
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(k_bvh_collapse, dim3((count + 255) / 256), dim3(256), 0, s, int(count),
                       work[cur], left, right, first, last, ids_sorted, tbox, nbox, h->nodes,
                       work[1 - cur], counters, counters + 1);
    e = hipMemcpyAsync(&count, counters, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cur = 1 - cur;
    ++depth;
  }
  unsigned nodes_used = 0;
  float rb[6];
  if (e == hipSuccess) e = hipMemcpyAsync(&nodes_used, counters + 1, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(rb, n > 1 ? nbox : tbox, 24, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return cleanup(e, "bvh build");
  h->num_nodes = nodes_used;
  h->depth = depth;
  // a node stacks at most 3 of its children
  h->stack_bound = depth > 1 ? 3u * (depth - 1u) : 0u;
  if (h->stack_bound > VRT_BVH_MAX_STACK) {
    const unsigned sb = h->stack_bound;
    (void)cleanup(hipSuccess, "bvh build");
    return fail(VRT_ERROR_UNSUPPORTED,
                "bvh: a traversal of this tree can stack %u entries, the kernels hold %d "
                "(4-wide depth %u)", sb, VRT_BVH_MAX_STACK, depth);
  }
  for (int a = 0; a < 3; ++a) {
    h->lo[a] = rb[a];
    h->hi[a] = rb[3 + a];
  }
  h->build_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = h;
  return VRT_OK;
}

}  // extern "C"
