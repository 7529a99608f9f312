// Device-side scene views and math shared by the gfx950 kernels.
//
// The host builds these from the vrt_* plain-old-data structs once per call;
// they travel to the GPU as kernel arguments (no per-frame allocation, unlike
// the reference's DevicePointer copies, DeviceArray.h:379-393).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vrt.h"

namespace vrt {

struct f3 {
  float x, y, z;
};
__host__ __device__ inline f3 mk(float a, float b, float c) { return f3{a, b, c}; }
__host__ __device__ inline f3 mk(float a) { return f3{a, a, a}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__host__ __device__ inline f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline float length(f3 a) { return sqrtf(dot(a, a)); }
__host__ __device__ inline f3 normalize(f3 a) { return a * (1.f / length(a)); }
__host__ __device__ inline float max3(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__host__ __device__ inline f3 fabs3(f3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__host__ __device__ inline float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// One volume instance, derived from vrt_volume on the host.
struct KVolume {
  int field_type;
  int dims[3];
  f3 blo, bhi;         // object-space bounds
  f3 S, T;             // voxel space = S * (p + T)
  float cell_size;
  const float *bricks; // structured: brick records of 9^3 voxels (VRT_BRICK_FLOATS each)
  uint32_t nbx, nby;   // structured: bricks per x row, per y column
  const float *leaves; // nanovdb: 512 values per leaf
  const float4 *tf;    // global copy of the table
  int tf_size;
  int tf_lds_offset;   // offset of this table in the block's LDS copy
  float vlo, vhi;
  float vrange_inv;    // 1 / (vhi - vlo), used when that range is a power of two
  int vrange_pow2;
  float unit_distance;
  const float *majorants;
  int gdims[3];
  const uint8_t *mbytes;  // majorant multiples of *mstep (optional)
  const float *mstep;
  int maj_lds_offset;     // byte offset of this grid in the workgroup's LDS, -1: global
  float A[9];          // affineInv
  f3 tinv;             // transInv
  uint32_t obj_id, inst_id;
  f3 glo, ghi;         // majorant grid bounds in voxel space
  // nanovdb: top table over 128^3 regions (>= 0 block index, < 0: -1 -
  // index into nv_consts), blocks of 16^3 slot entries (>= 0 leaf index
  // into `leaves`, else a constant), top origin (in 128-voxel units) / dims,
  // world -> index map, filter
  const int32_t *nv_top;
  const int32_t *nv_blocks;
  const float *nv_consts;
  const float *nv_slot_ranges;
  int nv_lo[3], nv_n[3];
  float nv_inv[9], nv_vec[3];
  int filter;
  // unstructured: point-location grid (list start per grid cell, dims
  // product + 1 entries; tetrahedron copies of four float4 {x, y, z, value}
  // in list order), its origin, cells per unit length and dims; per
  // tetrahedron {bbox min, value min}, {bbox max, value max}
  const uint32_t *ug_start;
  const float4 *ug_recs;
  const uint32_t *ug_cell;  // per list entry: primitive id (ANARI cell index)
  const float4 *ug_tets;
  uint32_t ug_num_tets;
  f3 ug_lo, ug_inv;
  int ug_dims[3];
};

struct KLight {
  int type;
  int visible;
  f3 dir;        // directional: unit vector towards the light
  f3 intensity;  // color * irradiance / intensity
  f3 v1, e1, e2; // quad: position, edges
  f3 nrm;        // quad: normalize(e1 x (e1 + e2))
  float a1, a2;  // quad: areas of its triangles {v1, e1, e1+e2}, {v1, e1+e2, e2}
};

// the scene's HDRI environment light (HDRI.cpp:102-162)
struct KHdri {
  const float4 *tex;   // w x h RGBA radiance, x fastest
  const float *rows;   // normalised CDF over the rows' totals (h entries)
  const float *cols;   // per row, the normalised CDF over its pixels (w entries each)
  int w, h;
  float scale;
  float to_world[9], to_local[9];  // row-major
};

struct KSurface {
  const float4 *nodes;  // 4-wide BVH nodes, 8 float4 each
  const float4 *tris;   // {v1, prim id bits}, e1, e2 per triangle
  f3 color;             // matte
  uint32_t obj_id, inst_id;
};

// A surface's BVH nodes staged in LDS: nodes [0, lds_nodes) -- breadth-
// first, so the top levels -- are read from the workgroup's copy at float4
// lds_first of the staged region (kept apart from KSurface, which shading
// copies per lane)
struct KStaged {
  int num_nodes, lds_nodes, lds_first, pad;
};

// An instanced mesh (vrt_instance) on the device
struct KInst {
  const float4 *nodes;  // the mesh's BVH (KSurface layout)
  const float4 *tris;
  float A[9];           // affineInv, row-major: object = A (world + tinv)
  float tinv[3];
  float N[9];           // normal transform transpose(A), row-major
  float color[3];       // matte
  uint32_t obj_id, inst_id;
  uint32_t index;       // position in the array given to vrt_tls_build
  uint32_t pad;
};

// The instance-level BVH: 4-wide nodes as KSurface's, whose leaves index
// `inst` (instances in leaf order)
struct KTls {
  const float4 *nodes;
  const KInst *inst;
  int num_inst;         // 0: no TLS
  int pad;
};

struct KScene {
  int num_volumes;
  int num_lights;
  int tf_lds_total;   // float4 entries of all tables staged in LDS
  int maj_lds_bytes;  // majorant bytes staged in LDS after the tables
  int num_surfaces;
  KVolume vol[VRT_MAX_VOLUMES];
  KLight light[VRT_MAX_LIGHTS];
  KSurface surf[VRT_MAX_SURFACES];
  KStaged staged[VRT_MAX_SURFACES];
  KTls tls;
  int hdri_light;  // index of the HDRI light, -1 for none
  KHdri hdri;
};

struct KCamera {
  f3 pos, U, V, W;
  float region[4];
  float shutter[2];
  // thin lens: lens radius / |U| and / |V| (the lens plane is spanned by U
  // and V; 0: pinhole), distance of the plane in focus along W
  float lens_u, lens_v;
  float focus_distance;
};

struct KParams {
  float4 bg;
  f3 ambient_color;
  float ambient_radiance;
  float occlusion_distance;
  int ambient_samples;
  int max_bounce;
  int pixel_samples;
  float sampling_rate_inv;
  uint32_t rng_offset;
  int num_clip_planes;
  float4 clip[VRT_MAX_CLIP_PLANES];
  float phase_g;
  int gradient_shading;
  int ratio_tracking;  // shadow / AO transmittance by ratio tracking
  // scheduling only (images do not change): a lane that continues its
  // bounce with shadow / AO rays traces them in the next passes while the
  // other lanes of its wave wait, and idle lanes are refilled between
  // bounces; 0: every lane traces one ray per pass, idle lanes refilled
  // before each pass
  int chain_passes;
  // scheduling only: idle lanes are refilled once at least this many lanes
  // of the wave are idle (or the wave has no active lane)
  int refill_min;
  // temporal anti-aliasing: previous and current view-projection (row-major
  // proj * view), blend weight
  int taa;
  float taa_alpha;
  float vp_prev[16], vp_curr[16];
};

struct KFrame {
  int width, height;
  int color_type;
  void *color;
  float4 *accum;
  float *depth;
  double *accum64;
  unsigned long long *counters;
  float *normal;  // g-buffer (float3 per pixel), optional
  float *albedo;
  uint32_t *prim_id, *obj_id, *inst_id;
  float4 *motion;                    // TAA: motion vector per pixel
  float4 *taa_curr, *taa_prev;       // TAA: current sample / history colour
  float *taa_curr_alb, *taa_prev_alb;  // TAA: albedo (float3) of each
};

// Per-frame buffers of the radiance-cache path tracer.  Paths that are not
// training paths end at their second volume vertex and leave a query (raw
// cache inputs + path throughput); training paths continue and leave one
// training sample (raw inputs + radiance of the path suffix / throughput).
struct KNrc {
  float4 *sample;        // per pixel: sample colour without the cache term
  int *query_of_pixel;   // per pixel: query index or -1
  float *q_raw;          // q_cap x 6
  float *q_T;            // q_cap x 3
  float *q_out;          // q_cap x 3 cache predictions
  float *t_raw;          // t_cap x 6
  float *t_target;       // t_cap x 3
  // self-training (textra > 0): per training sample, the query that ended
  // its path (-1: none) and that vertex's throughput over the query vertex's
  float *t_ratio;        // t_cap x 3
  int *t_query;          // t_cap
  uint32_t *ctr;         // [0] queries [1] training claims [2] vertex-2 paths [3] previous [2]
  int q_cap, t_cap;
  int train;             // record training paths this frame
  int qbounce;           // bounce index of the querying volume vertex (0 or 1)
  int textra;            // training paths end in a query this many volume vertices later (0: never)
  f3 pos_lo, pos_scale;  // world -> [0,1]^3 cache coordinates
};

}  // namespace vrt
