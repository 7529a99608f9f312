// vrt_bvh: triangle BVH for the surface hit path, replacing the reference's
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

  // collapse into 4-wide nodes, breadth first
  struct Slot { int node2; };
  std::vector<float4> nodes4;
  std::vector<float4> tris_out;
  tris_out.reserve(3 * n);
  std::vector<int> queue = {0};
  std::vector<int> node4_of;  // queue position -> node4 index (same order)
  for (size_t q = 0; q < queue.size(); ++q) {
    const Node2 &root = B.nodes[queue[q]];
    int kids[4];
    int nk = 0;
    if (root.count) {
      kids[nk++] = queue[q];  // a leaf root (tiny meshes): one leaf child
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
    float4 rec[8];
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
          e = int32_t(tris_out.size() / 3);
          cnt = uint32_t(c.count);
          for (int i = c.first; i < c.first + c.count; ++i) {
            const int t = B.order[i];
            const float *p = &tri[9 * size_t(t)];
            uint32_t pid = uint32_t(t);
            float pidf;
            std::memcpy(&pidf, &pid, 4);
            tris_out.push_back(make_float4(p[0], p[1], p[2], pidf));
            tris_out.push_back(make_float4(p[3], p[4], p[5], 0.f));
            tris_out.push_back(make_float4(p[6], p[7], p[8], 0.f));
          }
        } else {
          e = int32_t(queue.size());
          queue.push_back(kids[k]);
        }
      }
      std::memcpy(&lanes[6][k], &e, 4);
      std::memcpy(&lanes[7][k], &cnt, 4);
    }
    for (int r = 0; r < 8; ++r) rec[r] = make_float4(lanes[r][0], lanes[r][1], lanes[r][2], lanes[r][3]);
    nodes4.insert(nodes4.end(), rec, rec + 8);
  }

  // the traversal stacks, at a node, every inner child it hit but the one it
  // descends into: the most entries along any root-to-leaf path, bottom-up
  // (children follow their parent in the node array)
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
  const uint32_t stack_bound = nn ? need[0] : 0u;
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

}  // extern "C"
