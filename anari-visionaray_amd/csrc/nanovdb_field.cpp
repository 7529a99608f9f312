// vrt_nanovdb: conversion of a NanoVDB float grid buffer (format 32.x) into
// the GPU layout, replacing NanoVDBField::finalize's device copy of the
// buffer (NanoVDBField.cpp:30-65).
//
// Layout on the device, two levels instead of NanoVDB's four:
//  * a dense x-fastest top table over 128^3-voxel regions (the extent of a
//    lower internal node) covering the grid's bounds plus one region of
//    apron: a block index (>= 0) where a lower node exists, else -1 - c for
//    constant c (background, or the root / upper-node tile value there);
//  * blocks of 16^3 slot entries (one per 8^3 slot, x-fastest): a leaf index
//    (>= 0) or -1 - c for a lower-node tile;
//  * the leaf values, 512 floats per leaf in NanoVDB's in-leaf order
//    (i&7)<<6 | (j&7)<<3 | (k&7).
// A lookup is two table reads and one value read instead of a root tile
// search plus three node levels; grids with huge tiles stay small because a
// tile costs one top entry.
//
// Record offsets are those of the 32.7 format (GridData 672 B with the map at
// 296 and the world bbox at 560; TreeData 64 B; RootData 64 B + 32-B tiles;
// upper node table at 8256, lower at 1088, child masks after the 4096-B /
// 512-B value masks; leaf values at 96).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "nanovdb_internal.h"
#include "render_kernels.h"

using namespace vrt;

namespace {

template <class T> T rd(const uint8_t *p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

bool bit(const uint8_t *mask, uint32_t n) { return (mask[n >> 3] >> (n & 7)) & 1u; }

constexpr uint64_t kMaxTop = uint64_t(1) << 24;

struct Tables {
  int lo[3], n[3];
  std::vector<int32_t> top;
  std::vector<int32_t> blocks;
  std::unordered_map<uint32_t, int32_t> const_of_bits;
  std::vector<float> consts;

  int32_t constant(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    auto it = const_of_bits.find(b);
    if (it != const_of_bits.end()) return it->second;
    const int32_t c = int32_t(consts.size());
    consts.push_back(v);
    const_of_bits.emplace(b, c);
    return c;
  }
  // top-table index of the 128^3 region holding voxel o, or -1 outside
  int64_t top_index(const int o[3]) const {
    int t[3];
    for (int a = 0; a < 3; ++a) {
      t[a] = (o[a] >> 7) - lo[a];
      if (t[a] < 0 || t[a] >= n[a]) return -1;
    }
    return (int64_t(t[2]) * n[1] + t[1]) * n[0] + t[0];
  }
  // constant entry for every top region of the box [o, o + size)
  void fill_top(const int o[3], int64_t size, int32_t e) {
    int t0[3], t1[3];
    for (int a = 0; a < 3; ++a) {
      t0[a] = std::max(o[a] >> 7, lo[a]) - lo[a];
      const int64_t last = (int64_t(o[a]) + size - 1) >> 7;
      t1[a] = int(std::min<int64_t>(last, int64_t(lo[a]) + n[a] - 1)) - lo[a];
      if (t1[a] < t0[a]) return;
    }
    for (int z = t0[2]; z <= t1[2]; ++z)
      for (int y = t0[1]; y <= t1[1]; ++y)
        for (int x = t0[0]; x <= t1[0]; ++x) top[(size_t(z) * n[1] + y) * n[0] + x] = e;
  }
};

}  // namespace

namespace vrt {

void nanovdb_cell_voxels(const vrt_nanovdb_s &g, const float glo[3], const float ghi[3],
                         const int dims[3], int cx, int cy, int cz, int lo[3], int hi[3]) {
  const int c[3] = {cx, cy, cz};
  double wmin[3], wmax[3];
  for (int a = 0; a < 3; ++a) {
    const double w = (double(ghi[a]) - double(glo[a])) / double(dims[a]);
    wmin[a] = double(glo[a]) + double(c[a]) * w;
    wmax[a] = double(glo[a]) + double(c[a] + 1) * w;
  }
  double imin[3] = {INFINITY, INFINITY, INFINITY}, imax[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = 0; k < 8; ++k) {
    const double p[3] = {(k & 1 ? wmax[0] : wmin[0]) - g.vec_d[0],
                         (k & 2 ? wmax[1] : wmin[1]) - g.vec_d[1],
                         (k & 4 ? wmax[2] : wmin[2]) - g.vec_d[2]};
    for (int a = 0; a < 3; ++a) {
      const double i = g.inv_d[3 * a] * p[0] + g.inv_d[3 * a + 1] * p[1] + g.inv_d[3 * a + 2] * p[2];
      imin[a] = std::min(imin[a], i);
      imax[a] = std::max(imax[a], i);
    }
  }
  for (int a = 0; a < 3; ++a) {
    lo[a] = int(std::floor(imin[a])) - 1;
    hi[a] = int(std::floor(imax[a])) + 2;
  }
}

}  // namespace vrt

extern "C" {

int vrt_nanovdb_create(vrt_context ctx, const void *grid_buffer, uint64_t bytes,
                       vrt_nanovdb *out) {
  if (int r = check_ctx(ctx)) return r;
  if (!grid_buffer || !out) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const uint8_t *b = static_cast<const uint8_t *>(grid_buffer);
  if (bytes < 672 + 64 + 64) return fail(VRT_ERROR_INVALID_ARGUMENT, "buffer too small");
  const uint64_t magic = rd<uint64_t>(b);
  if (magic != 0x304244566f6e614eull && magic != 0x314244566f6e614eull)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "not a NanoVDB grid buffer");
  if ((rd<uint32_t>(b + 16) >> 21) != 32u)
    return fail(VRT_ERROR_UNSUPPORTED, "NanoVDB major version %u (32 supported)",
                rd<uint32_t>(b + 16) >> 21);
  if (rd<uint32_t>(b + 636) != 1u)
    return fail(VRT_ERROR_UNSUPPORTED, "NanoVDB grid type %u (float supported)",
                rd<uint32_t>(b + 636));
  const uint64_t grid_size = rd<uint64_t>(b + 32);
  if (grid_size > bytes) return fail(VRT_ERROR_INVALID_ARGUMENT, "grid size exceeds the buffer");

  vrt_nanovdb_s g;
  g.device = ctx->device;
  g.stream = ctx->stream;
  for (int i = 0; i < 9; ++i) g.inv_f[i] = rd<float>(b + 332 + 4 * i);
  for (int i = 0; i < 3; ++i) g.vec_f[i] = rd<float>(b + 368 + 4 * i);
  for (int i = 0; i < 9; ++i) g.inv_d[i] = rd<double>(b + 456 + 8 * i);
  for (int i = 0; i < 3; ++i) g.vec_d[i] = rd<double>(b + 528 + 8 * i);
  double wlo[3], whi[3];
  for (int a = 0; a < 3; ++a) {
    wlo[a] = rd<double>(b + 560 + 8 * a);
    whi[a] = rd<double>(b + 584 + 8 * a);
    // NanoVDBField::finalize casts the world bbox to float
    g.blo[a] = float(wlo[a]);
    g.bhi[a] = float(whi[a]);
  }
  const uint8_t *tree = b + 672;
  const uint8_t *root = tree + rd<int64_t>(tree + 24);
  if (root < b + 736 || size_t(root - b) + 64 > grid_size)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "root outside the grid");
  g.active = rd<uint64_t>(tree + 56);
  const uint32_t ntiles = rd<uint32_t>(root + 24);
  g.background = rd<float>(root + 28);
  if (size_t(root - b) + 64 + size_t(ntiles) * 32 > grid_size)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "root tiles exceed the grid");

  // top-table extent: the index image of the world bbox, taps one voxel
  // below and two above, one region of apron
  Tables T;
  if (!(wlo[0] <= whi[0] && wlo[1] <= whi[1] && wlo[2] <= whi[2])) {
    for (int a = 0; a < 3; ++a) {
      T.lo[a] = 0;
      T.n[a] = 1;
      g.blo[a] = g.bhi[a] = 0.f;
    }
  } else {
    double imin[3] = {INFINITY, INFINITY, INFINITY}, imax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 8; ++k) {
      const double p[3] = {(k & 1 ? whi[0] : wlo[0]) - g.vec_d[0],
                           (k & 2 ? whi[1] : wlo[1]) - g.vec_d[1],
                           (k & 4 ? whi[2] : wlo[2]) - g.vec_d[2]};
      for (int a = 0; a < 3; ++a) {
        const double i = g.inv_d[3 * a] * p[0] + g.inv_d[3 * a + 1] * p[1] + g.inv_d[3 * a + 2] * p[2];
        imin[a] = std::min(imin[a], i);
        imax[a] = std::max(imax[a], i);
      }
    }
    uint64_t total = 1;
    for (int a = 0; a < 3; ++a) {
      if (!(std::fabs(imin[a]) < 1e9 && std::fabs(imax[a]) < 1e9))
        return fail(VRT_ERROR_INVALID_ARGUMENT, "grid bounds out of range");
      const int v0 = int(std::floor(imin[a])) - 1, v1 = int(std::floor(imax[a])) + 2;
      T.lo[a] = (v0 >> 7) - 1;
      T.n[a] = (v1 >> 7) + 1 - T.lo[a] + 1;
      total *= uint64_t(T.n[a]);
    }
    if (total > kMaxTop)
      return fail(VRT_ERROR_UNSUPPORTED, "grid bounds need %llu top entries (limit %llu)",
                  (unsigned long long)total, (unsigned long long)kMaxTop);
  }
  T.consts.push_back(g.background);
  {
    uint32_t bb;
    std::memcpy(&bb, &g.background, 4);
    T.const_of_bits.emplace(bb, 0);
  }
  T.top.assign(size_t(T.n[0]) * T.n[1] * T.n[2], -1);

  // paint: root tiles, upper-node tiles, lower nodes (blocks), their tiles
  // and leaves
  std::vector<const uint8_t *> leaves;
  for (uint32_t t = 0; t < ntiles; ++t) {
    const uint8_t *tile = root + 64 + 32 * size_t(t);
    const uint64_t key = rd<uint64_t>(tile);
    const int o[3] = {int(uint32_t((key >> 42) & 0x1FFFFF) << 12),
                      int(uint32_t((key >> 21) & 0x1FFFFF) << 12), int(uint32_t(key & 0x1FFFFF) << 12)};
    const int64_t child = rd<int64_t>(tile + 8);
    if (child == 0) {
      T.fill_top(o, 4096, -1 - T.constant(rd<float>(tile + 20)));
      continue;
    }
    const uint8_t *up = root + child;
    if (up < b || size_t(up - b) + 270400 > grid_size)
      return fail(VRT_ERROR_INVALID_ARGUMENT, "upper node outside the grid");
    for (uint32_t nu = 0; nu < 32768; ++nu) {
      const int ou[3] = {o[0] + int(nu >> 10) * 128, o[1] + int((nu >> 5) & 31) * 128,
                         o[2] + int(nu & 31) * 128};
      const int64_t ti = T.top_index(ou);
      if (ti < 0) continue;  // outside the bounds (+ apron): never sampled
      const uint8_t *entry = up + 8256 + 8 * size_t(nu);
      if (!bit(up + 32 + 4096, nu)) {
        T.top[size_t(ti)] = -1 - T.constant(rd<float>(entry));
        continue;
      }
      const uint8_t *low = up + rd<int64_t>(entry);
      if (low < b || size_t(low - b) + 33856 > grid_size)
        return fail(VRT_ERROR_INVALID_ARGUMENT, "lower node outside the grid");
      const int32_t blk = int32_t(T.blocks.size() / 4096);
      T.top[size_t(ti)] = blk;
      T.blocks.resize(T.blocks.size() + 4096);
      int32_t *slots = T.blocks.data() + size_t(blk) * 4096;
      for (uint32_t nl = 0; nl < 4096; ++nl) {
        // NanoVDB child index x<<8 | y<<4 | z -> our x-fastest slot order
        const uint32_t sx = nl >> 8, sy = (nl >> 4) & 15, sz = nl & 15;
        int32_t &e = slots[(sz * 16 + sy) * 16 + sx];
        const uint8_t *le = low + 1088 + 8 * size_t(nl);
        if (!bit(low + 32 + 512, nl)) {
          e = -1 - T.constant(rd<float>(le));
          continue;
        }
        const uint8_t *leaf = low + rd<int64_t>(le);
        if (leaf < b || size_t(leaf - b) + 2144 > grid_size)
          return fail(VRT_ERROR_INVALID_ARGUMENT, "leaf outside the grid");
        e = int32_t(leaves.size());
        leaves.push_back(leaf);
      }
    }
  }
  if (leaves.size() > 0x7FFFFFFFull || T.blocks.size() / 4096 > 0x7FFFFFFFull)
    return fail(VRT_ERROR_UNSUPPORTED, "too many nodes");

  // upload
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = T.lo[a];
    g.n[a] = T.n[a];
  }
  g.leaf_count = leaves.size();
  g.block_count = T.blocks.size() / 4096;
  g.const_count = T.consts.size();
  std::vector<float> host_leaves(leaves.size() * 512);
  for (size_t i = 0; i < leaves.size(); ++i)
    std::memcpy(host_leaves.data() + 512 * i, leaves[i] + 96, 2048);
  auto *h = new vrt_nanovdb_s(g);
  auto cleanup = [&](hipError_t e, const char *what) {
    (void)hipFree(h->top);
    (void)hipFree(h->blocks);
    (void)hipFree(h->leaves);
    (void)hipFree(h->consts);
    (void)hipFree(h->slot_ranges);
    delete h;
    return hip_fail(e, what);
  };
  const size_t nblk = T.blocks.size();
  hipError_t e = hipMalloc(&h->top, T.top.size() * 4);
  if (e == hipSuccess) e = hipMalloc(&h->blocks, std::max<size_t>(1, nblk) * 4);
  if (e == hipSuccess) e = hipMalloc(&h->leaves, std::max<size_t>(1, host_leaves.size()) * 4);
  if (e == hipSuccess) e = hipMalloc(&h->consts, T.consts.size() * 4);
  if (e == hipSuccess) e = hipMalloc(&h->slot_ranges, std::max<size_t>(1, nblk) * 8);
  if (e != hipSuccess) return cleanup(e, "nanovdb allocation");
  e = hipMemcpyAsync(h->top, T.top.data(), T.top.size() * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && nblk)
    e = hipMemcpyAsync(h->blocks, T.blocks.data(), nblk * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && !host_leaves.empty())
    e = hipMemcpyAsync(h->leaves, host_leaves.data(), host_leaves.size() * 4,
                       hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h->consts, T.consts.data(), T.consts.size() * 4, hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess)
    e = launch_nanovdb_slot_ranges(h->blocks, nblk, h->leaves, h->consts, h->slot_ranges,
                                   ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // host copies go out of scope
  if (e != hipSuccess) return cleanup(e, "nanovdb upload");
  h->device_bytes = T.top.size() * 4 + nblk * 12 + host_leaves.size() * 4 + T.consts.size() * 4;
  *out = h;
  return VRT_OK;
}

int vrt_nanovdb_get_info(vrt_nanovdb g, vrt_nanovdb_info *info) {
  if (!g || !info) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  for (int a = 0; a < 3; ++a) {
    info->bounds_lo[a] = g->blo[a];
    info->bounds_hi[a] = g->bhi[a];
    info->top_origin[a] = g->lo[a];
    info->top_dims[a] = g->n[a];
  }
  info->background = g->background;
  info->leaves = g->leaf_count;
  info->blocks = g->block_count;
  info->constants = g->const_count;
  info->active_voxels = g->active;
  info->device_bytes = g->device_bytes;
  return VRT_OK;
}

// Frees the device copies; the creating context may already be gone, so
// only the device id is used (hipFree waits for work that reads them).
int vrt_nanovdb_destroy(vrt_nanovdb g) {
  if (!g) return VRT_OK;
  (void)hipSetDevice(g->device);
  (void)hipFree(g->top);
  (void)hipFree(g->blocks);
  (void)hipFree(g->leaves);
  (void)hipFree(g->consts);
  (void)hipFree(g->slot_ranges);
  delete g;
  return VRT_OK;
}

}  // extern "C"
