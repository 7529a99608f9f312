// C ABI (include/vrt.h): argument validation, host-side derivation of the
// kernel scene views, and launches on the context stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "abi_internal.h"
#include "bvh_internal.h"
#include "nanovdb_internal.h"
#include "nrc_internal.h"
#include "render_kernels.h"
#include "unstructured_internal.h"
#include "hdri_internal.h"
#include "vrt.h"

namespace vrt {

static thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(VRT_ERROR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int check_ctx(vrt_context ctx) {
  if (!ctx) return fail(VRT_ERROR_INVALID_ARGUMENT, "null context");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return VRT_OK;
}

}  // namespace vrt

namespace {

using namespace vrt;

// shared by both field types: transfer function, majorants, instance
int volume_common(const vrt_volume &in, int lds_offset, KVolume &v) {
  if (!in.tf_rgba || in.tf_size < 1)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "volume has no transfer function table");
  if (!(in.unit_distance > 0.f))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "unit_distance must be > 0");
  if (in.grid_dims[0] > 0 && (!in.majorants || in.grid_dims[1] < 1 || in.grid_dims[2] < 1))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "majorant grid incomplete");
  if (in.filter != VRT_FILTER_LINEAR && in.filter != VRT_FILTER_NEAREST)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "unknown filter %d", in.filter);
  v.tf = reinterpret_cast<const float4 *>(in.tf_rgba);
  v.tf_size = in.tf_size;
  v.tf_lds_offset = lds_offset;
  v.vlo = in.value_lo;
  v.vhi = in.value_hi;
  {
    // (val - lo) / (hi - lo) equals (val - lo) * (1 / (hi - lo)) bit for bit
    // when hi - lo is a power of two whose reciprocal is a normal float
    const float range = in.value_hi - in.value_lo;
    int e = 0;
    v.vrange_pow2 = range > 0.f && std::frexp(range, &e) == 0.5f && e > -120 && e < 120;
    v.vrange_inv = v.vrange_pow2 ? 1.f / range : 0.f;
  }
  v.unit_distance = in.unit_distance;
  v.majorants = in.majorants;
  for (int a = 0; a < 3; ++a) v.gdims[a] = in.grid_dims[0] > 0 ? in.grid_dims[a] : 0;
  v.mbytes = in.majorant_bytes;
  v.mstep = in.majorant_step;
  v.maj_lds_offset = -1;
  for (int i = 0; i < 9; ++i) v.A[i] = in.affine_inv[i];
  v.tinv = mk(in.trans_inv[0], in.trans_inv[1], in.trans_inv[2]);
  v.obj_id = in.object_id;
  v.inst_id = in.instance_id;
  v.filter = in.filter;
  return VRT_OK;
}

// NanoVDBField (NanoVDBField.cpp:30-65): bounds = the grid's world bbox as
// float, identity voxel-space transform, cell size 1; the majorant grid
// spans the bounds
int make_nanovdb_volume(const vrt_volume &in, int lds_offset, KVolume &v) {
  const vrt_nanovdb_s *g = in.nanovdb;
  if (!g) return fail(VRT_ERROR_INVALID_ARGUMENT, "nanovdb volume has no grid");
  v = KVolume{};
  v.field_type = VRT_FIELD_NANOVDB;
  v.blo = mk(g->blo[0], g->blo[1], g->blo[2]);
  v.bhi = mk(g->bhi[0], g->bhi[1], g->bhi[2]);
  v.S = mk(1.f);
  v.T = mk(0.f);
  v.glo = v.blo;
  v.ghi = v.bhi;
  v.cell_size = 1.f;
  v.leaves = g->leaves;
  v.nv_top = g->top;
  v.nv_blocks = g->blocks;
  v.nv_consts = g->consts;
  v.nv_slot_ranges = g->slot_ranges;
  for (int a = 0; a < 3; ++a) {
    v.nv_lo[a] = g->lo[a];
    v.nv_n[a] = g->n[a];
    v.nv_vec[a] = g->vec_f[a];
  }
  for (int i = 0; i < 9; ++i) v.nv_inv[i] = g->inv_f[i];
  return volume_common(in, lds_offset, v);
}

// UnstructuredField (UnstructuredField.cpp:280-285): identity voxel-space
// transform, bounds = union of the cells' boxes, cell size = the running
// average diagonal; the majorant grid spans the bounds
int make_unstructured_volume(const vrt_volume &in, int lds_offset, KVolume &v) {
  const vrt_unstructured_s *f = in.unstructured;
  if (!f) return fail(VRT_ERROR_INVALID_ARGUMENT, "unstructured volume has no field");
  v = KVolume{};
  v.field_type = VRT_FIELD_UNSTRUCTURED;
  v.blo = mk(f->blo[0], f->blo[1], f->blo[2]);
  v.bhi = mk(f->bhi[0], f->bhi[1], f->bhi[2]);
  v.S = mk(1.f);
  v.T = mk(0.f);
  v.glo = v.blo;
  v.ghi = v.bhi;
  v.cell_size = f->cell_size;
  v.ug_start = f->start;
  v.ug_recs = f->recs;
  v.ug_cell = f->rec_cell;
  v.ug_tets = f->tets;
  v.ug_num_tets = uint32_t(f->num_tets);
  v.ug_lo = mk(f->lo[0], f->lo[1], f->lo[2]);
  v.ug_inv = mk(f->inv_cell[0], f->inv_cell[1], f->inv_cell[2]);
  for (int a = 0; a < 3; ++a) v.ug_dims[a] = f->dims[a];
  return volume_common(in, lds_offset, v);
}

int make_volume(const vrt_volume &in, int lds_offset, KVolume &v) {
  if (in.field_type == VRT_FIELD_NANOVDB) return make_nanovdb_volume(in, lds_offset, v);
  if (in.field_type == VRT_FIELD_UNSTRUCTURED) return make_unstructured_volume(in, lds_offset, v);
  if (in.field_type != VRT_FIELD_STRUCTURED_REGULAR)
    return fail(VRT_ERROR_UNSUPPORTED, "field type %d not supported by this entry point",
                in.field_type);
  for (int a = 0; a < 3; ++a)
    if (in.dims[a] < 1 || in.dims[a] > 4096)
      return fail(VRT_ERROR_INVALID_ARGUMENT, "volume dims[%d]=%d out of range", a, in.dims[a]);
  if (!in.bricks) return fail(VRT_ERROR_INVALID_ARGUMENT, "volume has no brick records");
  v = KVolume{};
  v.field_type = in.field_type;
  for (int a = 0; a < 3; ++a) v.dims[a] = in.dims[a];
  v.blo = mk(in.bounds_lo[0], in.bounds_lo[1], in.bounds_lo[2]);
  v.bhi = mk(in.bounds_hi[0], in.bounds_hi[1], in.bounds_hi[2]);
  // voxelSpaceTransform = (scaling(1/bounds.size()), -bounds.min)
  // (StructuredRegularField.cpp:37-40)
  v.S = mk(1.f / (v.bhi.x - v.blo.x), 1.f / (v.bhi.y - v.blo.y), 1.f / (v.bhi.z - v.blo.z));
  v.T = -v.blo;
  v.cell_size = in.cell_size;
  v.bricks = in.bricks;
  v.nbx = uint32_t((in.dims[0] + 7) / 8);
  v.nby = uint32_t((in.dims[1] + 7) / 8);
  v.glo = mk(0.f);
  v.ghi = mk(1.f);
  return volume_common(in, lds_offset, v);
}

constexpr int kMaxLdsTf = 4096;           // float4 entries (64 KiB)
constexpr int kMaxLdsMajorants = 40960;   // bytes: a 32^3 grid (16^3-voxel cells at 512^3)

int make_scene(const vrt_scene *s, KScene &k) {
  if (!s) return fail(VRT_ERROR_INVALID_ARGUMENT, "null scene");
  if (s->num_volumes < 0 || s->num_volumes > VRT_MAX_VOLUMES)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "num_volumes %d out of range", s->num_volumes);
  if (s->num_lights < 0 || s->num_lights > VRT_MAX_LIGHTS)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "num_lights %d out of range", s->num_lights);
  k = KScene{};
  k.num_volumes = s->num_volumes;
  k.num_lights = s->num_lights;
  k.hdri_light = -1;
  int off = 0;
  for (int i = 0; i < s->num_volumes; ++i) {
    int r = make_volume(s->volumes[i], off, k.vol[i]);
    if (r) return r;
    off += s->volumes[i].tf_size;
  }
  if (off > kMaxLdsTf)
    return fail(VRT_ERROR_UNSUPPORTED, "transfer function tables exceed %d entries", kMaxLdsTf);
  k.tf_lds_total = off;
  // majorant grids with byte copies go to LDS while they fit; the others are
  // read as floats from global memory
  int mbytes = 0;
  for (int i = 0; i < s->num_volumes; ++i) {
    KVolume &v = k.vol[i];
    const int cells = v.gdims[0] * v.gdims[1] * v.gdims[2];
    if (cells > 0 && v.mbytes && v.mstep && mbytes + cells <= kMaxLdsMajorants) {
      v.maj_lds_offset = mbytes;
      mbytes += (cells + 15) & ~15;
    }
  }
  k.maj_lds_bytes = mbytes;
  for (int i = 0; i < s->num_lights; ++i) {
    const vrt_light &L = s->lights[i];
    KLight &o = k.light[i];
    o.type = L.type;
    o.visible = L.visible;
    if (L.type == VRT_LIGHT_DIRECTIONAL) {
      f3 d = mk(L.direction[0], L.direction[1], L.direction[2]);
      float len = length(d);
      if (!(len > 0.f)) return fail(VRT_ERROR_INVALID_ARGUMENT, "directional light with zero direction");
      // Directional.cpp:33 stores -direction: the sample direction points at the light
      o.dir = -(d * (1.f / len));
      o.intensity = mk(L.color[0], L.color[1], L.color[2]) * L.intensity;
    } else if (L.type == VRT_LIGHT_QUAD) {
      // QuadLight::finalize (Quad.cpp:30-52), tessellated as in
      // DeviceCopyableObjects.h:2191-2196
      o.v1 = mk(L.position[0], L.position[1], L.position[2]);
      o.e1 = mk(L.edge1[0], L.edge1[1], L.edge1[2]);
      o.e2 = mk(L.edge2[0], L.edge2[1], L.edge2[2]);
      if (!(length(o.e1) > 0.f) || !(length(o.e2) > 0.f))
        return fail(VRT_ERROR_INVALID_ARGUMENT, "quad light has zero-length edges");
      const f3 d = o.e1 + o.e2;
      o.nrm = normalize(cross(o.e1, d));
      o.a1 = 0.5f * length(cross(o.e1, d));
      o.a2 = 0.5f * length(cross(d, o.e2));
      o.intensity = mk(L.color[0], L.color[1], L.color[2]) * std::max(0.f, L.intensity);
    } else if (L.type == VRT_LIGHT_HDRI) {
      if (!L.hdri) return fail(VRT_ERROR_INVALID_ARGUMENT, "hdri light without a map");
      if (k.hdri_light >= 0) return fail(VRT_ERROR_UNSUPPORTED, "more than one hdri light");
      k.hdri_light = i;
      KHdri &H = k.hdri;
      H.tex = L.hdri->tex;
      H.rows = L.hdri->rows;
      H.cols = L.hdri->cols;
      H.w = L.hdri->w;
      H.h = L.hdri->h;
      H.scale = L.scale;
      if (!hdri_frame(L.up, L.direction, H.to_world, H.to_local))
        return fail(VRT_ERROR_INVALID_ARGUMENT, "hdri up and direction span no frame");
    } else {
      return fail(VRT_ERROR_UNSUPPORTED, "light type %d not supported by this entry point", L.type);
    }
  }
  if (s->num_surfaces < 0 || s->num_surfaces > VRT_MAX_SURFACES)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "num_surfaces %d out of range", s->num_surfaces);
  k.num_surfaces = s->num_surfaces;
  for (int i = 0; i < s->num_surfaces; ++i) {
    const vrt_surface &S = s->surfaces[i];
    if (!S.bvh) return fail(VRT_ERROR_INVALID_ARGUMENT, "surface %d has no bvh", i);
    k.surf[i].nodes = S.bvh->nodes;
    k.surf[i].tris = S.bvh->tris;
    k.staged[i].num_nodes = int(S.bvh->num_nodes);
    k.surf[i].color = mk(S.color[0], S.color[1], S.color[2]);
    k.surf[i].obj_id = S.object_id;
    k.surf[i].inst_id = S.instance_id;
  }
  if (s->tls) {
    k.tls.nodes = s->tls->nodes;
    k.tls.inst = s->tls->inst;
    k.tls.num_inst = int(s->tls->num_inst);
  }
  return VRT_OK;
}

int make_camera(const vrt_camera *c, KCamera &k) {
  if (!c) return fail(VRT_ERROR_INVALID_ARGUMENT, "null camera");
  f3 d = mk(c->direction[0], c->direction[1], c->direction[2]);
  f3 up = mk(c->up[0], c->up[1], c->up[2]);
  if (!(length(d) > 0.f) || !(length(up) > 0.f))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "camera direction/up must be non-zero");
  d = normalize(d);
  up = normalize(up);
  const float th = std::tan(c->fovy * 0.5f);
  f3 right = cross(d, up);
  if (!(length(right) > 0.f)) return fail(VRT_ERROR_INVALID_ARGUMENT, "camera up parallel to direction");
  right = normalize(right);
  const f3 upv = cross(right, d);
  k.pos = mk(c->position[0], c->position[1], c->position[2]);
  k.U = right * (th * c->aspect);
  k.V = upv * th;
  k.W = d;
  for (int i = 0; i < 4; ++i) k.region[i] = c->image_region[i];
  k.shutter[0] = c->shutter[0];
  k.shutter[1] = c->shutter[1];
  if (!(c->aperture_radius >= 0.f) || !std::isfinite(c->aperture_radius))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "camera aperture_radius must be finite and >= 0");
  if (c->aperture_radius > 0.f && !(c->focus_distance > 0.f && std::isfinite(c->focus_distance)))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "camera focus_distance must be finite and > 0");
  k.lens_u = c->aperture_radius > 0.f ? c->aperture_radius / length(k.U) : 0.f;
  k.lens_v = c->aperture_radius > 0.f ? c->aperture_radius / length(k.V) : 0.f;
  k.focus_distance = c->focus_distance;
  return VRT_OK;
}

int make_params(const vrt_params *p, KParams &k) {
  if (!p) return fail(VRT_ERROR_INVALID_ARGUMENT, "null params");
  if (p->pixel_samples < 1 || p->pixel_samples > 256)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "pixel_samples %d out of [1,256]", p->pixel_samples);
  if (p->max_bounce < 0 || p->max_bounce > 256)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "max_bounce %d out of [0,256]", p->max_bounce);
  if (p->ambient_samples < 0 || p->ambient_samples > 256)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "ambient_samples %d out of [0,256]", p->ambient_samples);
  k.bg = make_float4(p->background[0], p->background[1], p->background[2], p->background[3]);
  k.ambient_color = mk(p->ambient_color[0], p->ambient_color[1], p->ambient_color[2]);
  k.ambient_radiance = p->ambient_radiance;
  k.occlusion_distance = p->occlusion_distance;
  k.ambient_samples = p->ambient_samples;
  k.max_bounce = p->max_bounce;
  k.pixel_samples = p->pixel_samples;
  // Raycast.cpp:27 safe_rcp
  k.sampling_rate_inv = p->volume_sampling_rate > 0.f ? 1.f / p->volume_sampling_rate : 0.f;
  k.rng_offset = p->rng_offset;
  if (p->num_clip_planes < 0 || p->num_clip_planes > VRT_MAX_CLIP_PLANES)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "num_clip_planes %d out of [0,%d]", p->num_clip_planes,
                VRT_MAX_CLIP_PLANES);
  k.num_clip_planes = p->num_clip_planes;
  if (!(p->phase_g > -1.f && p->phase_g < 1.f))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "phase_g %g out of (-1, 1)", double(p->phase_g));
  k.phase_g = p->phase_g;
  k.gradient_shading = p->gradient_shading ? 1 : 0;
  k.ratio_tracking = p->ratio_tracking ? 1 : 0;
  if (p->pass_schedule < 0 || p->pass_schedule > 3)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "pass_schedule %d out of [0,3]", p->pass_schedule);
  set_pass_schedule(k, 2);  // the path tracer's entry points choose it for their scene
  k.taa = p->taa ? 1 : 0;
  k.taa_alpha = p->taa_alpha;
  for (int i = 0; i < VRT_MAX_CLIP_PLANES; ++i) {
    const float *c = p->clip_planes[i];
    k.clip[i] = i < p->num_clip_planes ? make_float4(c[0], c[1], c[2], c[3])
                                       : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  return VRT_OK;
}

int make_frame(const vrt_frame *f, KFrame &k) {
  if (!f) return fail(VRT_ERROR_INVALID_ARGUMENT, "null frame");
  if (f->width < 1 || f->height < 1 || f->width > 16384 || f->height > 16384)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "frame size %dx%d out of range", f->width, f->height);
  if (f->color_type != VRT_TYPE_FLOAT32_VEC4 && f->color_type != VRT_TYPE_UFIXED8_VEC4 &&
      f->color_type != VRT_TYPE_UFIXED8_RGBA_SRGB && f->color_type != 0)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "unsupported color type %d", f->color_type);
  if (f->color_type != 0 && !f->color)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "color channel requested without a buffer");
  k.width = f->width;
  k.height = f->height;
  k.color_type = f->color_type;
  k.color = f->color;
  k.accum = reinterpret_cast<float4 *>(f->accum);
  k.depth = f->depth;
  k.accum64 = f->accum64;
  k.counters = f->counters;
  k.motion = reinterpret_cast<float4 *>(f->motion_vec);
  k.taa_curr = reinterpret_cast<float4 *>(f->taa_curr);
  k.taa_prev = reinterpret_cast<float4 *>(f->taa_prev);
  k.taa_curr_alb = f->taa_curr_albedo;
  k.taa_prev_alb = f->taa_prev_albedo;
  k.normal = f->normal;
  k.albedo = f->albedo;
  k.prim_id = f->primitive_id;
  k.obj_id = f->object_id;
  k.inst_id = f->instance_id;
  if (!f->depth && (f->normal || f->albedo || f->primitive_id || f->object_id || f->instance_id))
    return fail(VRT_ERROR_INVALID_ARGUMENT,
                "g-buffer channels need the depth channel (Frame::fillGBuffer)");
  return VRT_OK;
}

}  // namespace

extern "C" {

const char *vrt_last_error(void) { return vrt::g_error.c_str(); }

int vrt_context_create(int device, vrt_context *out) {
  if (!out) return fail(VRT_ERROR_INVALID_ARGUMENT, "null output");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(VRT_ERROR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= n) return fail(VRT_ERROR_INVALID_ARGUMENT, "device %d of %d", device, n);
  VRT_HIP(hipSetDevice(device));
  auto *c = new vrt_context_s;
  c->device = device;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->t0);
  if (e == hipSuccess) e = hipEventCreate(&c->t1);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipMalloc(&c->queues, 2 * 8 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(c->queues, 0, 2 * 8 * sizeof(uint32_t));
  if (e == hipSuccess) e = vrt::static_queue_address(&c->static_queues);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "context creation");
  }
  *out = c;
  return VRT_OK;
}

int vrt_context_destroy(vrt_context ctx) {
  if (!ctx) return VRT_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipEventDestroy(ctx->t0);
  (void)hipEventDestroy(ctx->t1);
  (void)hipFree(ctx->queues);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return VRT_OK;
}

int vrt_context_stream(vrt_context ctx, void **s) {
  if (!ctx || !s) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  *s = ctx->stream;
  return VRT_OK;
}

int vrt_synchronize(vrt_context ctx) {
  if (int r = check_ctx(ctx)) return r;
  VRT_HIP(hipStreamSynchronize(ctx->stream));
  return VRT_OK;
}

int vrt_alloc(vrt_context ctx, size_t bytes, void **d) {
  if (int r = check_ctx(ctx)) return r;
  if (!d) return fail(VRT_ERROR_INVALID_ARGUMENT, "null output");
  *d = nullptr;
  if (bytes == 0) return VRT_OK;
  hipError_t e = hipMalloc(d, bytes);
  if (e == hipErrorOutOfMemory) return fail(VRT_ERROR_OUT_OF_MEMORY, "out of device memory (%zu bytes)", bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  return VRT_OK;
}

int vrt_free(vrt_context ctx, void *d) {
  if (int r = check_ctx(ctx)) return r;
  if (d) VRT_HIP(hipFree(d));
  return VRT_OK;
}

int vrt_upload(vrt_context ctx, void *dst, const void *src, size_t bytes) {
  if (int r = check_ctx(ctx)) return r;
  if (bytes == 0) return VRT_OK;
  if (!dst || !src) return fail(VRT_ERROR_INVALID_ARGUMENT, "null buffer");
  VRT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  VRT_HIP(hipStreamSynchronize(ctx->stream));
  return VRT_OK;
}

int vrt_download(vrt_context ctx, void *dst, const void *src, size_t bytes) {
  if (int r = check_ctx(ctx)) return r;
  if (bytes == 0) return VRT_OK;
  if (!dst || !src) return fail(VRT_ERROR_INVALID_ARGUMENT, "null buffer");
  VRT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  VRT_HIP(hipStreamSynchronize(ctx->stream));
  return VRT_OK;
}

int vrt_memset(vrt_context ctx, void *dst, int value, size_t bytes) {
  if (int r = check_ctx(ctx)) return r;
  if (bytes == 0) return VRT_OK;
  if (!dst) return fail(VRT_ERROR_INVALID_ARGUMENT, "null buffer");
  VRT_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
  return VRT_OK;
}

int vrt_device_copy(vrt_context ctx, void *dst, const void *src, size_t bytes) {
  if (int r = check_ctx(ctx)) return r;
  if (bytes == 0) return VRT_OK;
  if (!dst || !src) return fail(VRT_ERROR_INVALID_ARGUMENT, "null buffer");
  if ((bytes & 15) || (reinterpret_cast<uintptr_t>(dst) & 15) ||
      (reinterpret_cast<uintptr_t>(src) & 15))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "copy needs 16-byte aligned buffers and size");
  VRT_HIP(launch_copy(dst, src, bytes, ctx->stream));
  return VRT_OK;
}

int vrt_timer_start(vrt_context ctx) {
  if (int r = check_ctx(ctx)) return r;
  VRT_HIP(hipEventRecord(ctx->t0, ctx->stream));
  return VRT_OK;
}

int vrt_timer_stop(vrt_context ctx, float *ms) {
  if (int r = check_ctx(ctx)) return r;
  VRT_HIP(hipEventRecord(ctx->t1, ctx->stream));
  VRT_HIP(hipEventSynchronize(ctx->t1));
  float v = 0.f;
  VRT_HIP(hipEventElapsedTime(&v, ctx->t0, ctx->t1));
  if (ms) *ms = v;
  return VRT_OK;
}

size_t vrt_structured_brick_bytes(const int dims[3]) {
  if (!dims || dims[0] < 1 || dims[1] < 1 || dims[2] < 1) return 0;
  return size_t((dims[0] + 7) / 8) * ((dims[1] + 7) / 8) * ((dims[2] + 7) / 8) *
         VRT_BRICK_FLOATS * sizeof(float);
}

int vrt_structured_bricks(vrt_context ctx, const float *lin, const int dims[3], float *bricks) {
  if (int r = check_ctx(ctx)) return r;
  if (!lin || !bricks || !dims) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  for (int a = 0; a < 3; ++a)
    if (dims[a] < 1 || dims[a] > 4096)
      return fail(VRT_ERROR_INVALID_ARGUMENT, "dims[%d]=%d out of range", a, dims[a]);
  VRT_HIP(launch_bricks(lin, dims, bricks, ctx->stream));
  return VRT_OK;
}

int vrt_build_value_ranges(vrt_context ctx, const vrt_volume *vol, float *ranges) {
  if (int r = check_ctx(ctx)) return r;
  if (!vol || !ranges) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  KVolume v;
  vrt_volume tmp = *vol;
  if (tmp.grid_dims[0] < 1 || tmp.grid_dims[1] < 1 || tmp.grid_dims[2] < 1)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "grid_dims must be positive");
  if (!tmp.majorants) tmp.majorants = ranges;  // not read by this kernel
  if (!tmp.tf_rgba) {
    static const float dummy[4] = {0, 0, 0, 0};
    tmp.tf_rgba = dummy;
    tmp.tf_size = 1;
  }
  if (!(tmp.unit_distance > 0.f)) tmp.unit_distance = 1.f;
  if (int r = make_volume(tmp, 0, v)) return r;
  if (v.field_type == VRT_FIELD_NANOVDB) {
    // each cell's voxel box on the host (double), then a union of slot ranges
    const int *G = tmp.grid_dims;
    const size_t n = size_t(G[0]) * G[1] * G[2];
    std::vector<int32_t> boxes(6 * n);
    const float glo[3] = {v.glo.x, v.glo.y, v.glo.z}, ghi[3] = {v.ghi.x, v.ghi.y, v.ghi.z};
    for (int z = 0; z < G[2]; ++z)
      for (int y = 0; y < G[1]; ++y)
        for (int x = 0; x < G[0]; ++x)
          nanovdb_cell_voxels(*tmp.nanovdb, glo, ghi, G, x, y, z,
                              &boxes[6 * ((size_t(z) * G[1] + y) * G[0] + x)],
                              &boxes[6 * ((size_t(z) * G[1] + y) * G[0] + x) + 3]);
    int32_t *d = nullptr;
    VRT_HIP(hipMalloc(&d, boxes.size() * 4));
    hipError_t e = hipMemcpyAsync(d, boxes.data(), boxes.size() * 4, hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess)
      e = launch_nanovdb_value_ranges(v, d, ranges, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    VRT_HIP(e);
    return VRT_OK;
  }
  if (v.field_type == VRT_FIELD_UNSTRUCTURED) {
    VRT_HIP(launch_tet_value_ranges(v, ranges, ctx->stream));
    return VRT_OK;
  }
  VRT_HIP(launch_value_ranges(v, ranges, ctx->stream));
  return VRT_OK;
}

int vrt_tf_build(const float *colors, int num_colors, int channels, const float *opac,
                 int num_opac, const float uni[4], float *out) {
  if (!uni || (colors && (num_colors < 1 || (channels != 3 && channels != 4))) ||
      (opac && num_opac < 1))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad transfer function arrays");
  const size_t nc = colors ? size_t(num_colors) : 1, no = opac ? size_t(num_opac) : 1;
  const size_t n = std::max(nc, no);
  if (n > 0x7FFFFFFF) return fail(VRT_ERROR_UNSUPPORTED, "transfer function too large");
  if (!out) return int(n);
  // position of entry i on an array of m samples, and its lerp weight
  auto pos = [n](size_t i, size_t m, size_t &lo, size_t &hi, float &w) {
    const float p = n > 1 ? (float(i) / float(n - 1)) * float(m - 1) : 0.f;
    lo = size_t(std::floor(p));
    hi = size_t(std::ceil(p));
    w = p - std::floor(p);
  };
  for (size_t i = 0; i < n; ++i) {
    float a[4] = {uni[0], uni[1], uni[2], uni[3]}, b[4] = {uni[0], uni[1], uni[2], uni[3]};
    size_t lo, hi;
    float w;
    pos(i, nc, lo, hi, w);
    if (colors)
      for (int k = 0; k < channels; ++k) {
        a[k] = colors[lo * size_t(channels) + size_t(k)];
        b[k] = colors[hi * size_t(channels) + size_t(k)];
      }
    float c[4];
    for (int k = 0; k < 4; ++k) c[k] = a[k] + w * (b[k] - a[k]);
    if (opac) {
      pos(i, no, lo, hi, w);
      c[3] *= opac[lo] + w * (opac[hi] - opac[lo]);
    }
    for (int k = 0; k < 4; ++k) out[4 * i + size_t(k)] = c[k];
  }
  return int(n);
}

int vrt_compute_majorants(vrt_context ctx, const vrt_volume *vol, const float *ranges, float *maj,
                          uint8_t *bytes, float *step) {
  if (int r = check_ctx(ctx)) return r;
  if (!vol || !ranges || !maj) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  vrt_volume tmp = *vol;
  if (tmp.grid_dims[0] < 1 || tmp.grid_dims[1] < 1 || tmp.grid_dims[2] < 1)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "grid_dims must be positive");
  if (!tmp.majorants) tmp.majorants = maj;
  KVolume v;
  if (int r = make_volume(tmp, 0, v)) return r;
  VRT_HIP(launch_majorants(v, ranges, maj, bytes, step, ctx->stream));
  return VRT_OK;
}

int vrt_sample_field(vrt_context ctx, const vrt_volume *vol, int n, const float *pts,
                     float *values) {
  if (int r = check_ctx(ctx)) return r;
  if (!vol || n < 0 || (n > 0 && (!pts || !values)))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad sample_field arguments");
  if (n == 0) return VRT_OK;
  KVolume v;
  if (int r = make_volume(*vol, 0, v)) return r;
  VRT_HIP(launch_sample_field(v, n, pts, values, ctx->stream));
  return VRT_OK;
}

int vrt_intersect_surfaces(vrt_context ctx, const vrt_scene *scene, int n, const float *rays,
                           int any, float *t_out, int *prim_out, int *surf_out) {
  if (int r = check_ctx(ctx)) return r;
  if (n < 0 || (n > 0 && (!rays || !t_out || !prim_out)))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad intersect_surfaces arguments");
  KScene sc;
  if (int r = make_scene(scene, sc)) return r;
  if (n == 0) return VRT_OK;
  VRT_HIP(launch_intersect_surfaces(sc, n, rays, any, t_out, prim_out, surf_out, ctx->stream));
  return VRT_OK;
}

int vrt_delta_track(vrt_context ctx, const vrt_scene *scene, int n, const float *rays,
                    uint32_t *rng, float *t_out) {
  if (int r = check_ctx(ctx)) return r;
  if (n < 0 || (n > 0 && (!rays || !rng || !t_out)))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad delta_track arguments");
  KScene sc;
  if (int r = make_scene(scene, sc)) return r;
  if (sc.num_volumes < 1) return fail(VRT_ERROR_INVALID_ARGUMENT, "scene has no volume");
  if (n == 0) return VRT_OK;
  VRT_HIP(launch_delta_track(sc, n, rays, rng, t_out, ctx->stream));
  return VRT_OK;
}

int vrt_relmse(vrt_context ctx, const float *accum, const double *ref_sum, double ref_scale,
               int num_pixels, float eps, double *out) {
  if (int r = check_ctx(ctx)) return r;
  if (!accum || !ref_sum || !out || num_pixels < 1)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "bad relmse arguments");
  double *d = nullptr;
  VRT_HIP(hipMalloc(&d, sizeof(double)));
  hipError_t e = hipMemsetAsync(d, 0, sizeof(double), ctx->stream);
  if (e == hipSuccess)
    e = launch_relmse(reinterpret_cast<const float4 *>(accum), ref_sum, ref_scale, num_pixels,
                      eps, d, ctx->stream);
  double s = 0.0;
  if (e == hipSuccess) e = hipMemcpyAsync(&s, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "relmse");
  *out = s / (3.0 * double(num_pixels));
  return VRT_OK;
}

int vrt_frame_reset(vrt_context ctx, const vrt_frame *frame) {
  if (int r = check_ctx(ctx)) return r;
  KFrame F;
  if (int r = make_frame(frame, F)) return r;
  VRT_HIP(launch_frame_reset(F, ctx->stream));
  return VRT_OK;
}

int vrt_raycast(vrt_context ctx, const vrt_scene *scene, const vrt_camera *camera,
                const vrt_params *params, const vrt_frame *frame, int frame_id) {
  (void)frame_id;
  if (int r = check_ctx(ctx)) return r;
  KScene sc;
  KCamera cam;
  KParams P;
  KFrame F;
  if (int r = make_scene(scene, sc)) return r;
  if (int r = make_camera(camera, cam)) return r;
  if (int r = make_params(params, P)) return r;
  if (int r = make_frame(frame, F)) return r;
  // the fixed-step marcher reads no majorants
  sc.maj_lds_bytes = 0;
  for (int i = 0; i < sc.num_volumes; ++i) sc.vol[i].maj_lds_offset = -1;
  uint32_t *cur = ctx->queues + 8 * ctx->queue_parity;
  uint32_t *nxt = ctx->queues + 8 * (1 - ctx->queue_parity);
  VRT_HIP(launch_raycast(sc, cam, P, F, cur, nxt, ctx->num_cus, ctx->stream));
  ctx->queue_parity ^= 1;
  return VRT_OK;
}

// the camera's proj * view, row-major (pinhole_camera: perspective(fovy,
// aspect, .001, 1000) and the view of eye / dir / up, camera/Perspective.cpp:
// 36-40; visionaray's matrices, parity unpinned: gluPerspective, gluLookAt)
int camera_view_proj(const vrt_camera *c, float M[16]) {
  f3 f = mk(c->direction[0], c->direction[1], c->direction[2]);
  f3 up = mk(c->up[0], c->up[1], c->up[2]);
  const f3 eye = mk(c->position[0], c->position[1], c->position[2]);
  if (!(length(f) > 0.f) || !(length(up) > 0.f))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "camera direction/up must be non-zero");
  f = normalize(f);
  const f3 s = normalize(cross(f, up));
  const f3 u = cross(s, f);
  const float V[16] = {s.x, s.y, s.z, -dot(s, eye), u.x, u.y, u.z, -dot(u, eye),
                       -f.x, -f.y, -f.z, dot(f, eye), 0.f, 0.f, 0.f, 1.f};
  const float zn = .001f, zf = 1000.f;
  const float g = 1.f / std::tan(c->fovy * 0.5f);
  const float Pm[16] = {g / c->aspect, 0.f, 0.f, 0.f, 0.f, g, 0.f, 0.f,
                        0.f, 0.f, (zf + zn) / (zn - zf), (2.f * zf * zn) / (zn - zf),
                        0.f, 0.f, -1.f, 0.f};
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k) {
      float acc = 0.f;
      for (int j = 0; j < 4; ++j) acc += Pm[4 * r + j] * V[4 * j + k];
      M[4 * r + k] = acc;
    }
  return VRT_OK;
}

}  // extern "C"

namespace {

// vrt_pathtrace and vrt_pathtrace_noalloc: the same frame, the work-queue
// counters either the context's (alternating sets) or the library's
// file-scope array (cleared per call)
int pathtrace_frame(vrt_context ctx, const vrt_scene *scene, const vrt_camera *camera,
                    const vrt_params *params, const vrt_frame *frame, int accum_id,
                    bool static_queues) {
  if (int r = check_ctx(ctx)) return r;
  if (accum_id < 0) return fail(VRT_ERROR_INVALID_ARGUMENT, "accum_id must be >= 0");
  KScene sc;
  KCamera cam;
  KParams P;
  KFrame F;
  if (int r = make_scene(scene, sc)) return r;
  if (int r = make_camera(camera, cam)) return r;
  if (int r = make_params(params, P)) return r;
  if (int r = make_frame(frame, F)) return r;
  if (!F.accum) return fail(VRT_ERROR_INVALID_ARGUMENT, "pathtrace needs an accumulation buffer");
  if (P.taa) {
    if (!F.motion || !F.taa_curr || !F.taa_prev || !F.taa_curr_alb || !F.taa_prev_alb)
      return fail(VRT_ERROR_INVALID_ARGUMENT, "taa needs the frame's motion and taa buffers");
    if (int r = camera_view_proj(camera, P.vp_curr)) return r;
    if (params->has_prev_camera) {
      if (int r = camera_view_proj(&params->prev_camera, P.vp_prev)) return r;
    } else {
      std::copy(P.vp_curr, P.vp_curr + 16, P.vp_prev);
    }
  }
  set_pass_schedule(P, params->pass_schedule == 0 ? pathtrace_schedule(sc)
                                                  : params->pass_schedule);
  const size_t lds = pathtrace_lds_size(sc);
  const int per_cu = pathtrace_blocks_per_cu(sc, lds, false, P.gradient_shading != 0);
  plan_staged_nodes(sc, lds, per_cu);  // the top BVH nodes that fit beside it
  int blocks = std::max(1, ctx->num_cus) * per_cu;
  blocks = std::max(8, (blocks + 7) / 8 * 8);  // every XCD queue gets workgroups
  if (static_queues) {
    VRT_HIP(launch_pathtrace_static(sc, cam, P, F, accum_id, ctx->static_queues, blocks,
                                    ctx->stream));
  } else {
    uint32_t *cur = ctx->queues + 8 * ctx->queue_parity;
    uint32_t *nxt = ctx->queues + 8 * (1 - ctx->queue_parity);
    VRT_HIP(launch_pathtrace(sc, cam, P, F, accum_id, cur, nxt, blocks, ctx->stream));
    ctx->queue_parity ^= 1;
  }
  if (P.taa) VRT_HIP(launch_taa_resolve(F, P.taa_alpha, ctx->stream));
  return VRT_OK;
}

}  // namespace

extern "C" {

int vrt_pathtrace(vrt_context ctx, const vrt_scene *scene, const vrt_camera *camera,
                  const vrt_params *params, const vrt_frame *frame, int accum_id, int frame_id) {
  (void)frame_id;
  return pathtrace_frame(ctx, scene, camera, params, frame, accum_id, false);
}

int vrt_pathtrace_noalloc(vrt_context ctx, const vrt_scene *scene, const vrt_camera *camera,
                          const vrt_params *params, const vrt_frame *frame, int accum_id,
                          int frame_id) {
  (void)frame_id;
  if (int r = check_ctx(ctx)) return r;
  if (!frame) return fail(VRT_ERROR_INVALID_ARGUMENT, "null frame");
  if (int64_t(frame->width) * frame->height > VRT_NOALLOC_MAX_PIXELS)
    return fail(VRT_ERROR_INVALID_ARGUMENT,
                "vrt_pathtrace_noalloc: %d x %d pixels exceeds its maximum of %lld pixels "
                "(the range of its work-queue counters)",
                frame->width, frame->height, (long long)VRT_NOALLOC_MAX_PIXELS);
  if (!ctx->static_queues)
    return fail(VRT_ERROR_HIP, "vrt_pathtrace_noalloc: the library's queue counters are not "
                               "addressable on this device");
  return pathtrace_frame(ctx, scene, camera, params, frame, accum_id, true);
}

int vrt_pathtrace_nrc(vrt_context ctx, const vrt_scene *scene, const vrt_camera *camera,
                      const vrt_params *params, const vrt_frame *frame, vrt_nrc nrc,
                      const vrt_nrc_render *opts, int accum_id, int frame_id) {
  (void)frame_id;
  if (int r = check_ctx(ctx)) return r;
  if (!nrc || !opts) return fail(VRT_ERROR_INVALID_ARGUMENT, "null cache or options");
  if (nrc->ctx != ctx) return fail(VRT_ERROR_INVALID_ARGUMENT, "cache belongs to another context");
  if (accum_id < 0) return fail(VRT_ERROR_INVALID_ARGUMENT, "accum_id must be >= 0");
  if (opts->train_samples < 0 || opts->train_samples > nrc->cfg.max_batch * std::max(1, opts->train_steps))
    return fail(VRT_ERROR_INVALID_ARGUMENT, "train_samples %d exceeds max_batch x train_steps",
                opts->train_samples);
  if (opts->train && opts->train_steps < 1)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "train_steps must be >= 1 when training");
  if (opts->query_vertex < 0 || opts->query_vertex > 2)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "query_vertex must be 1 or 2 (0: 2)");
  if (opts->train_extra_vertices < 0 || opts->train_extra_vertices > 64)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "train_extra_vertices %d out of [0, 64]",
                opts->train_extra_vertices);
  const int qbounce = (opts->query_vertex == 0 ? 2 : opts->query_vertex) - 1;
  KScene sc;
  KCamera cam;
  KParams P;
  KFrame F;
  if (int r = make_scene(scene, sc)) return r;
  if (int r = make_camera(camera, cam)) return r;
  if (int r = make_params(params, P)) return r;
  if (P.taa) return fail(VRT_ERROR_UNSUPPORTED, "taa with the radiance cache is not supported");
  if (cam.lens_u > 0.f)
    return fail(VRT_ERROR_UNSUPPORTED,
                "a thin-lens camera (aperture_radius > 0) with the radiance cache is not supported");
  if (int r = make_frame(frame, F)) return r;
  if (!F.accum) return fail(VRT_ERROR_INVALID_ARGUMENT, "pathtrace needs an accumulation buffer");
  if (P.pixel_samples != 1)
    return fail(VRT_ERROR_UNSUPPORTED, "the radiance-cache path tracer takes one sample per pixel");
  if (qbounce == 0 && F.depth && (F.normal || F.albedo || F.prim_id || F.obj_id || F.inst_id))
    return fail(VRT_ERROR_UNSUPPORTED, "g-buffer channels with first-vertex cache queries");
  const size_t pixels = size_t(F.width) * F.height;
  if (int r = nrc_frame_buffers(nrc, pixels, std::max(1, opts->train_samples))) return r;
  KNrc N{};
  N.sample = nrc->sample;
  N.query_of_pixel = nrc->query_of_pixel;
  N.q_raw = nrc->q_raw;
  N.q_T = nrc->q_T;
  N.q_out = nrc->q_out;
  N.t_raw = nrc->t_raw;
  N.t_target = nrc->t_target;
  N.t_ratio = nrc->t_ratio;
  N.t_query = nrc->t_query;
  N.ctr = nrc->ctr;
  N.q_cap = int(pixels);
  N.t_cap = opts->train ? std::max(0, opts->train_samples) : 0;
  N.train = opts->train && opts->train_samples > 0;
  N.qbounce = qbounce;
  N.textra = opts->train_extra_vertices;
  N.pos_lo = mk(opts->bounds_lo[0], opts->bounds_lo[1], opts->bounds_lo[2]);
  for (int a = 0; a < 3; ++a)
    if (!(opts->bounds_hi[a] > opts->bounds_lo[a]))
      return fail(VRT_ERROR_INVALID_ARGUMENT, "cache bounds must have positive extent");
  N.pos_scale = mk(1.f / (opts->bounds_hi[0] - opts->bounds_lo[0]),
                   1.f / (opts->bounds_hi[1] - opts->bounds_lo[1]),
                   1.f / (opts->bounds_hi[2] - opts->bounds_lo[2]));

  set_pass_schedule(P, params->pass_schedule == 0 ? pathtrace_schedule(sc)
                                                  : params->pass_schedule);
  const size_t lds = pathtrace_lds_size(sc);
  const int per_cu = pathtrace_blocks_per_cu(sc, lds, true, P.gradient_shading != 0);
  plan_staged_nodes(sc, lds, per_cu);
  int blocks = std::max(1, ctx->num_cus) * per_cu;
  blocks = std::max(8, (blocks + 7) / 8 * 8);
  uint32_t *cur = ctx->queues + 8 * ctx->queue_parity;
  uint32_t *nxt = ctx->queues + 8 * (1 - ctx->queue_parity);
  VRT_HIP(launch_pathtrace(sc, cam, P, F, accum_id, cur, nxt, blocks, ctx->stream, &N));
  ctx->queue_parity ^= 1;
  // answer the queries (count taken on the GPU), add the cache term, write
  // the frame, then train on this frame's samples
  VRT_HIP(nrc_launch_infer(nrc->d, N.q_cap, N.q_raw, N.q_out, ctx->num_cus, ctx->stream,
                           &N.ctr[0]));
  VRT_HIP(launch_nrc_resolve(F, N, accum_id, ctx->stream));
  if (N.train && N.textra > 0) VRT_HIP(launch_nrc_train_targets(N, ctx->stream));
  if (N.train) {
    const int saved = nrc->t_cap;
    nrc->t_cap = N.t_cap;
    int r = nrc_train_frame(nrc, opts->train_steps);
    nrc->t_cap = saved;
    if (r) return r;
  }
  return VRT_OK;
}

}  // extern "C"
