// vrt_bvh: triangle BVH for the surface hit path (and vrt_tls, the instance
// BVH over meshes placed by transforms), replacing the reference's
// CPU-built visionaray BVH uploaded per geometry (DeviceBVH.h:153-255, 320;
// the reference's device traverses a binary BVH with a per-thread stack).
//
// Build (host): binary BVH by binned SAH (16 bins on the axis of largest
// centroid extent, leaves of <= 4 triangles, or <= 8 where no split helps),
// then collapsed into 4-wide nodes by repeatedly opening the child of largest
// surface area.  Layout on the device, one 128-byte node per 4 children so a
// wave reads a node with 8 coalesced 16-byte loads:
//   float4[0..5]  child boxes: lo.x, hi.x, lo.y, hi.y, lo.z, hi.z (lane c = child c)
//   float4[6]     child entry (int bits): node index, or first triangle of a leaf,
//                 -1 for an empty slot
//   float4[7]     triangle count (uint bits): 0 for an inner node
// Triangles are stored in leaf order as {v1, prim id bits}, e1, e2 (the
// visionaray basic_triangle form, edges formed in float from the vertices).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "abi_internal.h"
#include "bvh_internal.h"

using namespace vrt;

namespace {

struct Box {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void grow(const float *p) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  void grow(const Box &b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  float area() const {
    const float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    if (!(x >= 0.f && y >= 0.f && z >= 0.f)) return 0.f;
    return 2.f * (x * y + y * z + z * x);
  }
};

struct Node2 {
  Box box;
  int left = -1, right = -1;  // children (inner)
  int first = 0, count = 0;   // triangle range in `order` (leaf)
};

struct Builder {
  const std::vector<Box> &tri_box;
  const std::vector<float> &centre;  // 3 per triangle
  std::vector<int> order;
  std::vector<Node2> nodes;
  uint32_t depth = 0;

  Builder(const std::vector<Box> &b, const std::vector<float> &c) : tri_box(b), centre(c) {}

  int build(int first, int count, uint32_t level) {
    depth = std::max(depth, level + 1);
    const int id = int(nodes.size());
    nodes.emplace_back();
    Box box, cbox;
    for (int i = first; i < first + count; ++i) {
      box.grow(tri_box[order[i]]);
      cbox.grow(&centre[3 * size_t(order[i])]);
    }
    nodes[id].box = box;
    auto make_leaf = [&]() {
      nodes[id].first = first;
      nodes[id].count = count;
      return id;
    };
    if (count <= 4 || level >= 60) return make_leaf();
    int axis = 0;
    for (int a = 1; a < 3; ++a)
      if (cbox.hi[a] - cbox.lo[a] > cbox.hi[axis] - cbox.lo[axis]) axis = a;
    const float ext = cbox.hi[axis] - cbox.lo[axis];
    int mid = -1;
    if (ext > 0.f) {
      constexpr int kBins = 16;
      Box bb[kBins];
      int bn[kBins] = {0};
      const float scale = float(kBins) / ext;
      auto bin_of = [&](int t) {
        const int b = int((centre[3 * size_t(t) + axis] - cbox.lo[axis]) * scale);
        return std::min(std::max(b, 0), kBins - 1);
      };
      for (int i = first; i < first + count; ++i) {
        const int b = bin_of(order[i]);
        bb[b].grow(tri_box[order[i]]);
        ++bn[b];
      }
      float best = INFINITY;
      int best_split = -1;
      Box left[kBins];
      int ln[kBins];
      Box acc;
      int an = 0;
      for (int b = 0; b < kBins - 1; ++b) {
        acc.grow(bb[b]);
        an += bn[b];
        left[b] = acc;
        ln[b] = an;
      }
      acc = Box();
      an = 0;
      for (int b = kBins - 1; b > 0; --b) {
        acc.grow(bb[b]);
        an += bn[b];
        if (ln[b - 1] == 0 || an == 0) continue;
        const float cost = left[b - 1].area() * float(ln[b - 1]) + acc.area() * float(an);
        if (cost < best) {
          best = cost;
          best_split = b;
        }
      }
      const float leaf_cost = box.area() * float(count);
      if (best_split > 0 && (best < leaf_cost || count > 8)) {
        auto it = std::partition(order.begin() + first, order.begin() + first + count,
                                 [&](int t) { return bin_of(t) < best_split; });
        mid = int(it - order.begin());
      }
    }
    if (mid <= first || mid >= first + count) {
      if (count <= 8) return make_leaf();
      mid = first + count / 2;  // degenerate centroids: split the list
    }
    const int l = build(first, mid - first, level + 1);
    const int r = build(mid, first + count - mid, level + 1);
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
  }
};

// Collapses the binary tree into 4-wide nodes, breadth first, by
// repeatedly opening the child of largest surface area.  A leaf child's
// entry is the position of its first item in leaf order; `leaf_items`
// receives the items (indices into the builder's input) in that order.
void collapse4(const Builder &B, std::vector<float4> &nodes4, std::vector<int> &leaf_items) {
  std::vector<int> queue = {0};
  for (size_t q = 0; q < queue.size(); ++q) {
    const Node2 &root = B.nodes[queue[q]];
    int kids[4];
    int nk = 0;
    if (root.count) {
      kids[nk++] = queue[q];  // a leaf root (tiny inputs): one leaf child
    } else {
      kids[nk++] = root.left;
      kids[nk++] = root.right;
      while (nk < 4) {
        int best = -1;
        float ba = -1.f;
        for (int k = 0; k < nk; ++k) {
          const Node2 &c = B.nodes[kids[k]];
          if (!c.count && c.box.area() > ba) {
            ba = c.box.area();
            best = k;
          }
        }
        if (best < 0) break;
        const Node2 &c = B.nodes[kids[best]];
        kids[best] = c.left;
        kids[nk++] = c.right;
      }
    }
    float lanes[8][4];
    for (int k = 0; k < 4; ++k) {
      lanes[0][k] = lanes[2][k] = lanes[4][k] = 1.f;   // empty slot: inverted box
      lanes[1][k] = lanes[3][k] = lanes[5][k] = -1.f;
      int32_t e = -1;
      uint32_t cnt = 0;
      if (k < nk) {
        const Node2 &c = B.nodes[kids[k]];
        for (int a = 0; a < 3; ++a) {
          lanes[2 * a][k] = c.box.lo[a];
          lanes[2 * a + 1][k] = c.box.hi[a];
        }
        if (c.count) {
          e = int32_t(leaf_items.size());
          cnt = uint32_t(c.count);
          for (int i = c.first; i < c.first + c.count; ++i) leaf_items.push_back(B.order[i]);
        } else {
          e = int32_t(queue.size());
          queue.push_back(kids[k]);
        }
      }
      std::memcpy(&lanes[6][k], &e, 4);
      std::memcpy(&lanes[7][k], &cnt, 4);
    }
    for (int r = 0; r < 8; ++r)
      nodes4.push_back(make_float4(lanes[r][0], lanes[r][1], lanes[r][2], lanes[r][3]));
  }
}

// The traversal stacks, at a node, every inner child it hit but the one it
// descends into: the most entries along any root-to-leaf path, bottom-up
// (children follow their parent in the node array)
uint32_t stack_need(const std::vector<float4> &nodes4) {
  const size_t nn = nodes4.size() / 8;
  std::vector<uint32_t> need(nn, 0u);
  for (size_t i = nn; i-- > 0;) {
    const float4 ch = nodes4[8 * i + 6], cn = nodes4[8 * i + 7];
    const float chv[4] = {ch.x, ch.y, ch.z, ch.w}, cnv[4] = {cn.x, cn.y, cn.z, cn.w};
    uint32_t inner = 0, below = 0;
    for (int k = 0; k < 4; ++k) {
      int32_t c;
      uint32_t cnt;
      std::memcpy(&c, &chv[k], 4);
      std::memcpy(&cnt, &cnv[k], 4);
      if (c < 0 || cnt != 0u) continue;
      ++inner;
      below = std::max(below, need[size_t(c)]);
    }
    need[i] = (inner > 0 ? inner - 1 : 0) + below;
  }
  return nn ? need[0] : 0u;
}

}  // namespace

extern "C" {

int vrt_bvh_build(vrt_context ctx, const float *vertices, int64_t num_vertices,
                  const uint32_t *indices, int64_t num_triangles, vrt_bvh *out) {
  if (int r = check_ctx(ctx)) return r;
  if (!out || !vertices || num_triangles < 1 || num_vertices < 3)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad bvh arguments");
  if (num_triangles > (int64_t(1) << 30)) return fail(VRT_ERROR_UNSUPPORTED, "too many triangles");
  if (!indices && num_vertices < 3 * num_triangles)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "triangle soup needs 3 vertices per triangle");
  *out = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  const size_t n = size_t(num_triangles);
  std::vector<float> tri(9 * n);  // v1, e1, e2
  std::vector<Box> boxes(n);
  std::vector<float> centre(3 * n);
  for (size_t i = 0; i < n; ++i) {
    uint32_t idx[3] = {uint32_t(3 * i), uint32_t(3 * i + 1), uint32_t(3 * i + 2)};
    if (indices)
      for (int k = 0; k < 3; ++k) {
        idx[k] = indices[3 * i + k];
        if (int64_t(idx[k]) >= num_vertices)
          return fail(VRT_ERROR_INVALID_ARGUMENT, "triangle %zu index out of range", i);
      }
    const float *v[3] = {vertices + 3 * size_t(idx[0]), vertices + 3 * size_t(idx[1]),
                         vertices + 3 * size_t(idx[2])};
    for (int a = 0; a < 3; ++a) {
      tri[9 * i + a] = v[0][a];
      tri[9 * i + 3 + a] = v[1][a] - v[0][a];
      tri[9 * i + 6 + a] = v[2][a] - v[0][a];
    }
    for (int k = 0; k < 3; ++k) boxes[i].grow(v[k]);
    for (int a = 0; a < 3; ++a) centre[3 * i + a] = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
  }
  Builder B(boxes, centre);
  B.order.resize(n);
  for (size_t i = 0; i < n; ++i) B.order[i] = int(i);
  B.nodes.reserve(2 * n / 4 + 16);
  B.build(0, int(n), 0);

  std::vector<float4> nodes4;
  std::vector<int> leaf_items;
  collapse4(B, nodes4, leaf_items);
  std::vector<float4> tris_out;
  tris_out.reserve(3 * n);
  for (const int t : leaf_items) {
    const float *p = &tri[9 * size_t(t)];
    uint32_t pid = uint32_t(t);
    float pidf;
    std::memcpy(&pidf, &pid, 4);
    tris_out.push_back(make_float4(p[0], p[1], p[2], pidf));
    tris_out.push_back(make_float4(p[3], p[4], p[5], 0.f));
    tris_out.push_back(make_float4(p[6], p[7], p[8], 0.f));
  }
  const uint32_t stack_bound = stack_need(nodes4);
  if (stack_bound > VRT_BVH_MAX_STACK)
    return fail(VRT_ERROR_UNSUPPORTED,
                "bvh: a traversal of this tree can stack %u entries, the kernels hold %d "
                "(binary depth %u)", stack_bound, VRT_BVH_MAX_STACK, B.depth);

  auto *h = new vrt_bvh_s;
  h->device = ctx->device;
  h->stack_bound = stack_bound;
  h->num_nodes = uint32_t(nodes4.size() / 8);
  h->num_tris = uint32_t(n);
  h->depth = B.depth;
  for (int a = 0; a < 3; ++a) {
    h->lo[a] = B.nodes[0].box.lo[a];
    h->hi[a] = B.nodes[0].box.hi[a];
  }
  hipError_t e = hipMalloc(&h->nodes, nodes4.size() * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&h->tris, tris_out.size() * sizeof(float4));
  if (e == hipSuccess)
    e = hipMemcpy(h->nodes, nodes4.data(), nodes4.size() * sizeof(float4), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->tris, tris_out.data(), tris_out.size() * sizeof(float4), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(h->nodes);
    (void)hipFree(h->tris);
    delete h;
    return hip_fail(e, "bvh upload");
  }
  h->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = h;
  return VRT_OK;
}

int vrt_bvh_get_info(vrt_bvh b, vrt_bvh_info *info) {
  if (!b || !info) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  info->num_nodes = b->num_nodes;
  info->num_triangles = b->num_tris;
  info->depth = b->depth;
  for (int a = 0; a < 3; ++a) {
    info->bounds_lo[a] = b->lo[a];
    info->bounds_hi[a] = b->hi[a];
  }
  info->build_ms = b->build_ms;
  info->device_bytes = uint64_t(b->num_nodes) * 128 + uint64_t(b->num_tris) * 48;
  info->stack_bound = b->stack_bound;
  return VRT_OK;
}

int vrt_bvh_destroy(vrt_bvh b) {
  if (!b) return VRT_OK;
  (void)hipSetDevice(b->device);
  (void)hipFree(b->nodes);
  (void)hipFree(b->tris);
  delete b;
  return VRT_OK;
}

int vrt_instance_transforms(const float xf[12], float affine_inv[9], float trans_inv[3],
                            float normal_xfm[9]) {
  if (!xf || !affine_inv || !trans_inv || !normal_xfm)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "null transform argument");
  double L[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) L[3 * r + c] = double(xf[4 * r + c]);
  for (int i = 0; i < 12; ++i)
    if (!std::isfinite(xf[i])) return fail(VRT_ERROR_INVALID_ARGUMENT, "non-finite instance transform");
  // inverse(top_left(xfm)) by cofactors (Instance.cpp:94)
  const double det = L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) +
                     L[2] * (L[3] * L[7] - L[4] * L[6]);
  if (!(det != 0.0) || !std::isfinite(det))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "singular instance transform");
  const double inv[9] = {(L[4] * L[8] - L[5] * L[7]) / det, (L[2] * L[7] - L[1] * L[8]) / det,
                         (L[1] * L[5] - L[2] * L[4]) / det, (L[5] * L[6] - L[3] * L[8]) / det,
                         (L[0] * L[8] - L[2] * L[6]) / det, (L[2] * L[3] - L[0] * L[5]) / det,
                         (L[3] * L[7] - L[4] * L[6]) / det, (L[1] * L[6] - L[0] * L[7]) / det,
                         (L[0] * L[4] - L[1] * L[3]) / det};
  for (int i = 0; i < 9; ++i) {
    affine_inv[i] = float(inv[i]);
    if (!std::isfinite(affine_inv[i]))
      return fail(VRT_ERROR_INVALID_ARGUMENT, "instance transform not invertible in float");
  }
  for (int k = 0; k < 3; ++k) trans_inv[k] = -xf[4 * k + 3];  // Instance.cpp:95
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) normal_xfm[3 * r + c] = affine_inv[3 * c + r];  // :96
  return VRT_OK;
}

int vrt_tls_build(vrt_context ctx, const vrt_instance *ins, int n, vrt_tls *out) {
  if (int r = check_ctx(ctx)) return r;
  if (!out || !ins || n < 1) return fail(VRT_ERROR_INVALID_ARGUMENT, "bad tls arguments");
  if (n > (1 << 24)) return fail(VRT_ERROR_UNSUPPORTED, "too many instances (%d)", n);
  *out = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<Box> boxes(static_cast<size_t>(n));
  std::vector<float> centre(3 * size_t(n));
  std::vector<KInst> recs(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    const vrt_instance &I = ins[i];
    if (!I.bvh) return fail(VRT_ERROR_INVALID_ARGUMENT, "instance %d has no bvh", i);
    if (I.bvh->device != ctx->device)
      return fail(VRT_ERROR_INVALID_ARGUMENT, "instance %d: bvh on another device", i);
    KInst &k = recs[size_t(i)];
    std::memset(&k, 0, sizeof(k));
    k.nodes = I.bvh->nodes;
    k.tris = I.bvh->tris;
    if (vrt_instance_transforms(I.transform, k.A, k.tinv, k.N) != VRT_OK)
      return fail(VRT_ERROR_INVALID_ARGUMENT, "instance %d: singular or non-finite transform", i);
    for (int a = 0; a < 3; ++a) k.color[a] = I.color[a];
    k.obj_id = I.object_id;
    k.inst_id = I.instance_id;
    k.index = uint32_t(i);
    // world box: the mesh box's corners under [L | t] in double, rounded outwards
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 8; ++c) {
      const double p[3] = {(c & 1) ? I.bvh->hi[0] : I.bvh->lo[0],
                           (c & 2) ? I.bvh->hi[1] : I.bvh->lo[1],
                           (c & 4) ? I.bvh->hi[2] : I.bvh->lo[2]};
      for (int r = 0; r < 3; ++r) {
        const double w = double(I.transform[4 * r]) * p[0] + double(I.transform[4 * r + 1]) * p[1] +
                         double(I.transform[4 * r + 2]) * p[2] + double(I.transform[4 * r + 3]);
        lo[r] = std::min(lo[r], w);
        hi[r] = std::max(hi[r], w);
      }
    }
    for (int a = 0; a < 3; ++a) {
      boxes[size_t(i)].lo[a] = std::nextafter(float(lo[a]), -INFINITY);
      boxes[size_t(i)].hi[a] = std::nextafter(float(hi[a]), INFINITY);
      centre[3 * size_t(i) + a] = 0.5f * (boxes[size_t(i)].lo[a] + boxes[size_t(i)].hi[a]);
    }
  }
  Builder B(boxes, centre);
  B.order.resize(size_t(n));
  for (int i = 0; i < n; ++i) B.order[size_t(i)] = i;
  B.build(0, n, 0);
  std::vector<float4> nodes4;
  std::vector<int> leaf_items;
  collapse4(B, nodes4, leaf_items);
  const uint32_t stack_bound = stack_need(nodes4);
  if (stack_bound > VRT_BVH_MAX_STACK)
    return fail(VRT_ERROR_UNSUPPORTED, "tls: a walk of this tree can stack %u entries, the "
                "kernels hold %d", stack_bound, VRT_BVH_MAX_STACK);
  std::vector<KInst> leaf_recs;
  leaf_recs.reserve(size_t(n));
  for (const int it : leaf_items) leaf_recs.push_back(recs[size_t(it)]);

  auto *h = new vrt_tls_s;
  h->device = ctx->device;
  h->num_nodes = uint32_t(nodes4.size() / 8);
  h->num_inst = uint32_t(n);
  h->stack_bound = stack_bound;
  for (int a = 0; a < 3; ++a) {
    h->lo[a] = B.nodes[0].box.lo[a];
    h->hi[a] = B.nodes[0].box.hi[a];
  }
  hipError_t e = hipMalloc(&h->nodes, nodes4.size() * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&h->inst, leaf_recs.size() * sizeof(KInst));
  if (e == hipSuccess)
    e = hipMemcpy(h->nodes, nodes4.data(), nodes4.size() * sizeof(float4), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->inst, leaf_recs.data(), leaf_recs.size() * sizeof(KInst),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(h->nodes);
    (void)hipFree(h->inst);
    delete h;
    return hip_fail(e, "tls upload");
  }
  h->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = h;
  return VRT_OK;
}

int vrt_tls_get_info(vrt_tls t, vrt_tls_info *info) {
  if (!t || !info) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  info->num_instances = t->num_inst;
  info->num_nodes = t->num_nodes;
  info->stack_bound = t->stack_bound;
  for (int a = 0; a < 3; ++a) {
    info->bounds_lo[a] = t->lo[a];
    info->bounds_hi[a] = t->hi[a];
  }
  info->build_ms = t->build_ms;
  return VRT_OK;
}

int vrt_tls_destroy(vrt_tls t) {
  if (!t) return VRT_OK;
  (void)hipSetDevice(t->device);
  (void)hipFree(t->nodes);
  (void)hipFree(t->inst);
  delete t;
  return VRT_OK;
}

}  // extern "C"
