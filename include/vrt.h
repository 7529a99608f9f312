/*
 * vrt.h -- C ABI of the MI355X (gfx950) per-pixel rendering kernels.
 *
 * This is the operator-level boundary under the ANARI device: it replaces
 *   VisionarayRendererRaycast::renderFrame   (renderer/Raycast_impl.h:14-21)
 *   VisionarayRendererPathtrace::renderFrame (renderer/Pathtrace_impl.h:15-22)
 * whose arguments were DevicePointer<DeviceObjectRegistry/RendererState/
 * dco::Frame/dco::Camera>, uint2 size, SyncContext, worldID, frameID.  Here
 * every argument is a plain-old-data struct passed by value to the kernel
 * (no per-frame device allocation), all buffers are device pointers owned by
 * the caller, and all work is enqueued on the context's stream.
 *
 * Conventions
 *  - Every entry point returns VRT_OK (0) or a negative vrt_status; nothing
 *    throws across the ABI.  vrt_last_error() gives a message.
 *  - Buffers passed as device pointers are never freed by the library.
 *  - A frame is not re-entrant: one render call per frame at a time.
 */
#ifndef VRT_H
#define VRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VRT_API __attribute__((visibility("default")))

typedef enum vrt_status {
  VRT_OK = 0,
  VRT_ERROR_INVALID_ARGUMENT = -1,
  VRT_ERROR_HIP = -2,
  VRT_ERROR_OUT_OF_MEMORY = -3,
  VRT_ERROR_UNSUPPORTED = -4,
  VRT_ERROR_NO_DEVICE = -5
} vrt_status;

/* ---------------------------------------------------------------------- */
/* Scene description (plain old data; pointers are device pointers)       */
/* ---------------------------------------------------------------------- */

#define VRT_FIELD_STRUCTURED_REGULAR 0
#define VRT_FIELD_NANOVDB 1
#define VRT_FIELD_UNSTRUCTURED 2

#define VRT_FILTER_LINEAR 0
#define VRT_FILTER_NEAREST 1

/* A transferFunction1D volume over one spatial field, placed by one
 * instance.  Mirrors dco::Volume + dco::SpatialField + dco::GridAccel +
 * dco::TransferFunction1D (DeviceCopyableObjects.h:331-638) and the
 * instance's affineInv/transInv (DeviceCopyableObjects.h:1659-1680). */
typedef struct vrt_volume {
  int field_type;      /* VRT_FIELD_* */
  int dims[3];         /* voxels per axis */
  float bounds_lo[3];  /* object-space bounds: origin ..            */
  float bounds_hi[3];  /*   origin + (dims-1)*spacing               */
  float cell_size;     /* min(spacing) (StructuredRegularField.cpp:40) */
  /* structuredRegular: the brick records of vrt_structured_bricks */
  const float *bricks;
  /* nanovdb: the converted grid (vrt_nanovdb_create, replacing the device
     copy of the grid buffer, NanoVDBField.cpp:29-64); bounds and cell size
     then come from the grid and dims / bricks are unused */
  struct vrt_nanovdb_s *nanovdb;
  int filter;          /* VRT_FILTER_*: nanovdb "filter" (NanoVDBField.cpp:27) */
  /* unstructured: the tetrahedral field (vrt_unstructured_create); bounds
     and cell size then come from the field, voxel space = world space */
  struct vrt_unstructured_s *unstructured;
  /* transfer function: tf_size float4 entries, value range, unitDistance */
  const float *tf_rgba;
  int tf_size;
  float value_lo, value_hi;
  float unit_distance;
  /* majorant grid over voxel space [0,1]^3 (grid_dims[0] == 0: none, the
     whole box is one segment with majorant 1).  Majorants are multiples
     q * step of step = (max TF opacity)/255 rounded up, q in 0..255;
     majorant_bytes holds q per cell and *majorant_step the step (device;
     optional -- without them the kernels read the float values) */
  const float *majorants;
  int grid_dims[3];
  const uint8_t *majorant_bytes;
  const float *majorant_step;
  /* instance: object = affine_inv * (world + trans_inv) */
  float affine_inv[9]; /* row-major 3x3 */
  float trans_inv[3];
  /* ids reported in the frame's objectId / instanceId channels: the
     volume's and its instance's ANARI "id" (Volume.cpp:25, Instance "id";
     ~0u when unset) */
  uint32_t object_id;
  uint32_t instance_id;
} vrt_volume;

#define VRT_LIGHT_DIRECTIONAL 0
#define VRT_LIGHT_QUAD 1
#define VRT_LIGHT_HDRI 2

/* An HDRI environment map on the device (vrt_hdri_create). */
typedef struct vrt_hdri_s *vrt_hdri;

typedef struct vrt_light {
  int type;
  int visible;
  float direction[3]; /* directional: ANARI "direction" (light travel);
                         hdri: "direction" (HDRI.cpp:106, default (0,0,1)) */
  float position[3];  /* quad */
  float edge1[3];
  float edge2[3];
  float color[3];
  float intensity;    /* directional: irradiance; quad: intensity */
  /* hdri (HDRI.cpp:102-162; at most one per scene): the map, "up"
     (default (1,0,0)) and "scale" (default 1) */
  vrt_hdri hdri;
  float up[3];
  float scale;
} vrt_light;

#define VRT_MAX_VOLUMES 4
#define VRT_MAX_LIGHTS 8
#define VRT_MAX_CLIP_PLANES 8

/* Triangle BVH (vrt_bvh_build) of one mesh; see below. */
typedef struct vrt_bvh_s *vrt_bvh;

/* A triangle surface with a matte material (Triangle.cpp + Matte.cpp:
 * color; the reference's other material and attribute paths are not part of
 * this build), identity instance transform. */
typedef struct vrt_surface {
  vrt_bvh bvh;
  float color[3];
  uint32_t object_id;    /* Surface "id" (Surface.cpp:21), ~0u when unset */
  uint32_t instance_id;  /* the instance's "id", ~0u when unset */
} vrt_surface;

#define VRT_MAX_SURFACES 4

/* A triangle mesh placed by an instance transform (dco::Instance,
 * DeviceCopyableObjects.h:1659-1680, with the matte surface above):
 * world = L * object + t for the row-major 3x4 `transform` [L | t].  Rays
 * meet it in object space, ori' = affineInv (ori + transInv), dir' =
 * affineInv dir (intersect(Ray, Instance), :1772-1814), so hit distances
 * stay world distances; normals go back by transpose(affineInv)
 * (Instance.cpp:88-97, Pathtrace_impl.cpp:243-247). */
typedef struct vrt_instance {
  vrt_bvh bvh;
  float transform[12];
  float color[3];
  uint32_t object_id;    /* Surface "id" (Surface.cpp:21), ~0u when unset */
  uint32_t instance_id;  /* the instance's "id", ~0u when unset */
} vrt_instance;

/* Instance-level BVH over any number of instances (the world's TLS over
 * its instances, scene/VisionarayScene.cpp:69-117, World.cpp:188-244):
 * a 4-wide BVH over the instances' world boxes, built on the host and
 * uploaded with the instance records (vrt_tls_build).  A scene's surface
 * hit is the closest over its surfaces and its TLS's instances; a surface
 * keeps ties with an instance, and among instances the lower index in the
 * array given to vrt_tls_build wins ties. */
typedef struct vrt_tls_s *vrt_tls;

typedef struct vrt_scene {
  int num_volumes;
  vrt_volume volumes[VRT_MAX_VOLUMES];
  int num_lights;
  vrt_light lights[VRT_MAX_LIGHTS];
  int num_surfaces;
  vrt_surface surfaces[VRT_MAX_SURFACES];
  vrt_tls tls;  /* instanced meshes (NULL: none) */
} vrt_scene;

/* ANARI "perspective" camera (camera/Perspective.cpp:14-44,
 * camera/Camera.cpp:32-40). */
typedef struct vrt_camera {
  float position[3];
  float direction[3];
  float up[3];
  float fovy;   /* radians */
  float aspect;
  float image_region[4]; /* lo.x lo.y hi.x hi.y */
  float shutter[2];
  /* thin lens (camera/Perspective.cpp:21-22, 42-43: apertureRadius,
   * focusDistance): the lens radius (0: pinhole, no random numbers drawn)
   * and the distance of the plane in focus along the view direction (> 0
   * when the radius is) */
  float aperture_radius;
  float focus_distance;
} vrt_camera;

/* RendererState subset (renderer/common.h:55-91) with the ANARI renderer
 * parameter defaults of Renderer.cpp:42-53, Pathtrace.cpp:19-25,
 * Raycast.cpp:19. */
typedef struct vrt_params {
  float background[4];       /* (0,0,0,1) */
  float ambient_color[3];    /* (1,1,1) */
  float ambient_radiance;    /* 0.2 */
  float occlusion_distance;  /* 1e20 */
  int ambient_samples;       /* 1 */
  int max_bounce;            /* 7 */
  int pixel_samples;         /* 1 */
  float volume_sampling_rate;/* 0.5 (raycast) */
  /* extension: pixel streams are seeded with (pixelID, accumID + rng_offset);
     0 reproduces the reference (Pathtrace_impl.cpp:512), other values give
     accumulations that are independent of it */
  uint32_t rng_offset;
  /* clip planes (rendererState.clipPlanes, Renderer.cpp:44-66): each ray keeps
     the half-space N.x <= D of every plane (N, D) = clip_planes[i]
     (clipRay, renderer/common.h:1469-1482) */
  int num_clip_planes;
  float clip_planes[VRT_MAX_CLIP_PLANES][4];
  /* extension: Henyey-Greenstein anisotropy g in (-1, 1) of the volume phase
     function.  0 is the reference's isotropic phase (uniform sphere sample,
     NEE without phase factor, Pathtrace_impl.cpp:330-339) exactly; other
     values sample HG around the ray direction with the same two random
     numbers and weight volume NEE by 4 pi p_HG, which is 1 at g = 0. */
  float phase_g;
  /* gradientShading (Renderer.cpp:50, default 0): the ray caster shades
     each sample as matte under every light along the field's central-
     difference gradient (VolumeIntegration.h:420-459); the path tracer takes
     that gradient as the volume normal for ambient occlusion and its NEE as
     matte (Pathtrace_impl.cpp:204-212, 311-328) */
  int gradient_shading;
  /* path tracer "taa" / "taaAlpha" (Pathtrace.cpp:24-25, default 0 / 0.3):
     the frame pass records the raw sample, its albedo and motion vector
     (frame taa buffers below) and a resolve pass blends it with the history
     at the motion-compensated position (Frame.cpp:339-381,
     DeviceCopyableObjects.h:2693-2712).  Motion vectors project the hits
     with the previous and the current camera (prev_camera; has_prev_camera
     0 on the first frame: no motion). */
  int taa;
  float taa_alpha;
  int has_prev_camera;
  vrt_camera prev_camera;
  /* extension (default 0): 1 = shadow and ambient-occlusion rays estimate the
     volumes' transmittance by ratio tracking -- the same free-flight steps
     and majorant cells as delta tracking, each tentative collision
     multiplying the estimate by 1 - sigma/majorant -- instead of the
     reference's binary visibility (DeviceCopyableObjects.h:722-757,
     Pathtrace_impl.cpp:441-444, 461-466).  Same expectation, lower
     variance; the pixels' random streams then differ from the reference's. */
  int ratio_tracking;
  /* extension, path-tracer scheduling only (images are identical for
     every value): 1 = a lane's shadow and ambient-occlusion passes follow
     its radiance pass back to back while the rest of its wave waits, idle
     lanes are refilled between bounces; 2 = one ray per lane and pass, idle
     lanes refilled before every pass; 3 = 2 with idle lanes refilled only
     16 at a time; 0 (default) = chosen per scene from measurements: 3 with
     a NanoVDB field, 2 with a tetrahedral field, 1 otherwise. */
  int pass_schedule;
} vrt_params;

/* ANARI data types used for frame channels (ANARI 1.0 enum values). */
#define VRT_TYPE_FLOAT32 1068
#define VRT_TYPE_FLOAT32_VEC4 1071
#define VRT_TYPE_UFIXED8_VEC4 1039
#define VRT_TYPE_UFIXED8_RGBA_SRGB 2003

/* dco::Frame (DeviceCopyableObjects.h:2615-2661) */
typedef struct vrt_frame {
  int width, height;
  int color_type;     /* VRT_TYPE_* */
  void *color;        /* width*height pixels of color_type */
  float *accum;       /* float4 running mean (stochastic renderers) */
  float *depth;       /* optional float depth channel (may be NULL) */
  double *accum64;    /* optional float4-as-double running sum (FP64
                         long-run reference accumulation; may be NULL) */
  unsigned long long *counters; /* optional: [0] field lookups (trilinear
                                   taps of one sample, or one point
                                   location), [1] path-tracer samples whose
                                   camera ray hit the scene, [2] unused */
  /* optional g-buffer channels (Frame::fillGBuffer,
     DeviceCopyableObjects.h:2745-2774): the first radiance hit of the
     pixel's closest sample -- geometric normal (float3), albedo (float3),
     primitive / object / instance id (uint32, ~0u for none).  As in the
     reference they are written only together with `depth`, and only where
     the sample's depth is <= the stored depth (which the accumulation reset
     sets to 1e31). */
  float *normal;
  float *albedo;
  uint32_t *primitive_id;
  uint32_t *object_id;
  uint32_t *instance_id;
  /* temporal anti-aliasing (vrt_params.taa): per pixel the motion vector
     (float4), the current and previous (history) colour (float4) and
     albedo (float3); required when taa is set */
  float *motion_vec;
  float *taa_curr;
  float *taa_prev;
  float *taa_curr_albedo;
  float *taa_prev_albedo;
} vrt_frame;

/* ---------------------------------------------------------------------- */
/* Context, memory                                                         */
/* ---------------------------------------------------------------------- */
typedef struct vrt_context_s *vrt_context;

VRT_API const char *vrt_last_error(void);
VRT_API int vrt_context_create(int device, vrt_context *out);
VRT_API int vrt_context_destroy(vrt_context ctx);
VRT_API int vrt_context_stream(vrt_context ctx, void **hip_stream);
VRT_API int vrt_synchronize(vrt_context ctx);

VRT_API int vrt_alloc(vrt_context ctx, size_t bytes, void **dptr);
VRT_API int vrt_free(vrt_context ctx, void *dptr);
VRT_API int vrt_upload(vrt_context ctx, void *dst, const void *src, size_t bytes);
VRT_API int vrt_download(vrt_context ctx, void *dst, const void *src, size_t bytes);
VRT_API int vrt_memset(vrt_context ctx, void *dst, int value, size_t bytes);

/* Device-to-device copy by a plain streaming kernel (16 B per lane), also
 * the calibration kernel for HBM byte counters. */
VRT_API int vrt_device_copy(vrt_context ctx, void *dst, const void *src, size_t bytes);

/* Event timing on the context stream: record a start/stop pair around
 * enqueued work and read the elapsed milliseconds. */
VRT_API int vrt_timer_start(vrt_context ctx);
VRT_API int vrt_timer_stop(vrt_context ctx, float *ms);

/* ---------------------------------------------------------------------- */
/* Spatial field + majorant grid                                           */
/* ---------------------------------------------------------------------- */

/* The renderers' layout of a structuredRegular field: bricks of 8^3
 * voxel cells, brick (bx, by, bz) x-fastest, each a record of
 * VRT_BRICK_FLOATS fp32 values holding the 9^3 voxels 8bx .. 8bx+8 (and so
 * on per axis; x fastest, then y, then z; indices past the last voxel
 * repeat it, which is clamp addressing) and 7 pad values (records are
 * 128-byte multiples).  Every trilinear footprint then lies in one record:
 * a lookup forms one address and reads its 8 taps at fixed offsets (0, 1,
 * 9, 10, 81, 82, 90, 91 values), for 1.42x the voxel memory.
 * Bytes needed for a dims[0]*dims[1]*dims[2] field (0 for bad dims). */
#define VRT_BRICK_FLOATS 736
VRT_API size_t vrt_structured_brick_bytes(const int dims[3]);

/* x-fastest linear fp32 voxels (device) -> brick records (device).
 * Replaces the hipArray texture upload of StructuredRegularField::finalize
 * (StructuredRegularField.cpp:42-76). */
VRT_API int vrt_structured_bricks(vrt_context ctx, const float *linear, const int dims[3],
                                  float *bricks);

/* NanoVDB float grid (NanoVDBField.cpp:25-65).  The host buffer of one
 * float grid in the NanoVDB 32.x format (the "data" array of a "nanovdb"
 * spatial field) is converted once into the GPU layout: a dense top table of
 * 128^3-voxel regions over the grid's bounds (+1 region apron), each a
 * constant (background or tile value) or a block of 16^3 slot entries (leaf
 * index or lower-node tile constant), and the leaf values in NanoVDB's
 * in-leaf order.  Lookups inside the top table reproduce
 * NanoGrid<float>::getValue exactly; outside it they read the background.
 * The sampler is the one NanoVDBField uses: createSampler<0|1> at
 * worldToIndexF(P) (DeviceCopyableObjects.h:554-569). */
typedef struct vrt_nanovdb_s *vrt_nanovdb;

typedef struct vrt_nanovdb_info {
  float bounds_lo[3], bounds_hi[3]; /* GridData::mWorldBBox, as float */
  float background;
  int top_origin[3];                /* index of top entry (0,0,0) / 128 */
  int top_dims[3];                  /* top table dims */
  uint64_t blocks;                  /* lower nodes inside the bounds */
  uint64_t leaves;                  /* leaves copied to the device */
  uint64_t constants;               /* distinct constants (background first) */
  uint64_t active_voxels;           /* TreeData::mVoxelCount */
  uint64_t device_bytes;            /* tables + leaves + constants + slot ranges */
} vrt_nanovdb_info;

VRT_API int vrt_nanovdb_create(vrt_context ctx, const void *grid_buffer, uint64_t bytes,
                               vrt_nanovdb *out);
VRT_API int vrt_nanovdb_get_info(vrt_nanovdb grid, vrt_nanovdb_info *info);
VRT_API int vrt_nanovdb_destroy(vrt_nanovdb grid);

/* Triangle BVH of a mesh given as host arrays: num_vertices float3 positions
 * and num_triangles uint3 indices (NULL: a triangle soup of 3 vertices per
 * triangle).  Built on the host by binned SAH, collapsed to 4-wide nodes and
 * uploaded; replaces the reference's per-geometry BVH build and upload
 * (DeviceBVH.h:153-255, 320; Triangle.cpp). */
/* Entries of the kernels' per-ray traversal stack.  A built tree whose
 * traversal could stack more (stack_bound below) is rejected with
 * VRT_ERROR_UNSUPPORTED rather than traversed with dropped subtrees. */
#define VRT_BVH_MAX_STACK 64

typedef struct vrt_bvh_info {
  uint32_t num_nodes, num_triangles, depth;
  float bounds_lo[3], bounds_hi[3];
  double build_ms;
  uint64_t device_bytes;
  /* most entries the traversal stack can hold for this tree (host build:
     the exact maximum over root-to-leaf paths of the inner children each
     node stacks; GPU build: 3 per 4-wide level below the root) */
  uint32_t stack_bound;
} vrt_bvh_info;

VRT_API int vrt_bvh_build(vrt_context ctx, const float *vertices, int64_t num_vertices,
                          const uint32_t *indices, int64_t num_triangles, vrt_bvh *out);
/* The same BVH built on the GPU from device arrays (vertices: float3,
 * indices: uint3 or NULL for a soup), for meshes rebuilt at every commit:
 * a linear BVH (Morton order, Karras 2012) collapsed into the same 4-wide
 * layout, leaves of <= 4 triangles.  Hits are identical to the host build's
 * (closest hit, ties to the lower triangle index); the tree differs.
 * Replaces the reference's host build + upload (DeviceBVH.h:153-255, 320). */
VRT_API int vrt_bvh_build_gpu(vrt_context ctx, const float *vertices, int64_t num_vertices,
                              const uint32_t *indices, int64_t num_triangles, vrt_bvh *out);
VRT_API int vrt_bvh_get_info(vrt_bvh bvh, vrt_bvh_info *info);
VRT_API int vrt_bvh_destroy(vrt_bvh bvh);

/* The instance's ray and normal transforms (Instance::finalize,
 * Instance.cpp:88-97): affine_inv = inverse(L) (row-major; evaluated in
 * double by cofactors, rounded to float), trans_inv = -t, normal_xfm =
 * transpose(affine_inv).  VRT_ERROR_INVALID_ARGUMENT for a singular or
 * non-finite L.  vrt_tls_build uses exactly these values. */
VRT_API int vrt_instance_transforms(const float transform[12], float affine_inv[9],
                                    float trans_inv[3], float normal_xfm[9]);

typedef struct vrt_tls_info {
  uint32_t num_instances, num_nodes;
  /* most entries the instance walk can stack (<= VRT_BVH_MAX_STACK) */
  uint32_t stack_bound;
  float bounds_lo[3], bounds_hi[3];  /* world box of all instances */
  double build_ms;
} vrt_tls_info;

/* Builds the TLS of n >= 1 instances (host binned SAH over their world
 * boxes, the corners of each mesh box under its transform, rounded
 * outwards).  The meshes' BVHs must outlive the TLS.  Replaces the world's
 * instance BVH build and upload (VisionarayScene.cpp:69-117). */
VRT_API int vrt_tls_build(vrt_context ctx, const vrt_instance *instances, int n, vrt_tls *out);
VRT_API int vrt_tls_get_info(vrt_tls tls, vrt_tls_info *info);
VRT_API int vrt_tls_destroy(vrt_tls tls);

/* HDRI environment light (HDRI::finalize, HDRI.cpp:31-162): `radiance` is
   width x height RGB float32, x fastest.  Builds the per-row and last-column
   CDFs of the max-component luminance (makeCDF) on the host and uploads them
   with an RGBA copy of the map.  Light samples pick a cell by the CDFs and
   take the direction of its corner (sampleCDF.h, DeviceCopyableObjects.h:
   2398-2418); lookups are bilinear with clamp addressing. */
VRT_API int vrt_hdri_create(vrt_context ctx, const float *radiance, int width, int height,
                            vrt_hdri *out);
VRT_API int vrt_hdri_destroy(vrt_hdri hdri);


/* Unstructured (tetrahedral) field: UnstructuredField::commitParameters /
 * finalize (UnstructuredField.cpp:17-198).  The host arrays are the field's
 * ANARI parameters: "vertex.position" (float3 x num_vertices), "vertex.data"
 * (float x num_vertices, or NULL), "index" (uint32 x num_indices),
 * "cell.index" (uint32 x num_cells: first index of each cell), "cell.type"
 * (uint8 x num_cells: 10 = VTK tetrahedron, 0 = its barney code) and
 * "cell.data" (float x num_cells, or NULL).  vertex.data or cell.data must be
 * given.  A cell's value at a point inside it is the barycentric blend of its
 * vertex data (intersectTet, Plane.h:56-96); when its first vertex has no
 * data (NaN, or no vertex.data) it is the cell's cell.data value (0 without
 * one) (DeviceCopyableObjects.h:128-164).  Cells with fewer than 4 indices
 * or a type/count mismatch are skipped, as the reference skips them; cells
 * of other types (pyramid, wedge, hexahedron) are rejected with
 * VRT_ERROR_UNSUPPORTED, as are indices outside the vertex array.
 *
 * Point location replaces the element BVH's any-hit query
 * (DeviceCopyableObjects.h:482-527) with a uniform grid over the bounds
 * (about one grid cell per tetrahedron) whose cells list a copy of every
 * tetrahedron overlapping them, most-covering first: a lookup tests the
 * listed tetrahedra in order and takes the first that contains the point.
 * Bounds = the union of the cells' bounding boxes (the BVH root,
 * UnstructuredField.cpp:295-332); cell_size = the reference's running
 * "avgCellDiagonal", which divides each cell's diagonal by the number of
 * cells kept so far (UnstructuredField.cpp:120-170, 281). */
typedef struct vrt_unstructured_s *vrt_unstructured;

typedef struct vrt_unstructured_info {
  float bounds_lo[3], bounds_hi[3];
  float cell_size;
  uint64_t num_cells;       /* tetrahedra kept */
  uint64_t skipped_cells;   /* cells skipped as invalid */
  int grid_dims[3];         /* point-location grid */
  uint64_t list_entries;    /* tetrahedron copies over all grid cells */
  uint32_t max_list;
  double mean_list;
  double expected_tests;    /* estimated tetrahedron tests per lookup */
  double build_ms;          /* host build of the grid */
  uint64_t device_bytes;
} vrt_unstructured_info;

VRT_API int vrt_unstructured_create(vrt_context ctx, const float *vertex_position,
                                    const float *vertex_data, int64_t num_vertices,
                                    const uint32_t *index, int64_t num_indices,
                                    const uint32_t *cell_index, const uint8_t *cell_type,
                                    const float *cell_data, int64_t num_cells,
                                    vrt_unstructured *out);
VRT_API int vrt_unstructured_get_info(vrt_unstructured field, vrt_unstructured_info *info);
VRT_API int vrt_unstructured_destroy(vrt_unstructured field);

/* Per-macrocell value ranges (float2 min/max) of a field over a
 * grid_dims grid in voxel space, conservative for trilinear sampling.
 * Replaces StructuredRegularField::buildGrid + rasterizeBox
 * (StructuredRegularField.cpp:95-127, GridAccel.h:13-67), which the
 * reference skips on GPU builds. */
VRT_API int vrt_build_value_ranges(vrt_context ctx, const vrt_volume *vol,
                                   float *ranges);

/* The transferFunction1D table (TransferFunction1D::finalize,
 * TransferFunction1D.cpp:66-111), on the host: max(num_colors,
 * num_opacities) float4 entries (1 when neither array is given); entry i
 * lerps the color array (color_channels 3 or 4; NULL: uniform_rgba) at
 * i/(n-1)*(num_colors-1), with alpha uniform_rgba[3] for 3-channel colors,
 * and multiplies alpha by the opacity array lerped the same way.
 * uniform_rgba = {color.xyz, opacity * color.w} of the volume's uniform
 * "color"/"opacity" parameters.  Returns the entry count and writes the
 * entries when out_rgba is not NULL; a negative status on bad arguments. */
VRT_API int vrt_tf_build(const float *colors, int num_colors, int color_channels,
                         const float *opacities, int num_opacities,
                         const float uniform_rgba[4], float *out_rgba);

/* Majorants (max TF opacity over each macrocell's value range, rounded up
 * to a multiple of the 8-bit step, see vrt_volume).  Replaces
 * GridAccel::computeMaxOpacities (GridAccel.cpp:69-107).  `bytes` (one per
 * cell) and `step` (one float) may be NULL. */
VRT_API int vrt_compute_majorants(vrt_context ctx, const vrt_volume *vol,
                                  const float *ranges, float *majorants,
                                  uint8_t *bytes, float *step);

/* Field values at n voxel-space points (device float3 array -> device
 * float array): sampleField (DeviceCopyableObjects.h:474-570) -- trilinear
 * on the brick records, the NanoVDB sampler, or the containing tetrahedron's
 * value; a point outside every tetrahedron (sampleField returns false) gives
 * NaN. */
VRT_API int vrt_sample_field(vrt_context ctx, const vrt_volume *vol, int n,
                             const float *points, float *values);

/* Delta tracking of n world-space rays through scene volume 0:
 * rays = n x {ori.xyz, dir.xyz, tmin, tmax} (device), rng = n LCG states
 * (device, advanced in place), t_out = hit distance or +inf (device).
 * The intersect(Ray, Volume) collision sampler of
 * DeviceCopyableObjects.h:680-796. */
/* Closest (any = 0) or any (any = 1) surface hit of n rays {ori xyz, dir
 * xyz, tmin, tmax} (device float[8] each): t (+inf on a miss), the
 * triangle's index in its mesh (-1) and, when surf_out is not NULL, what
 * was hit: surface i as i, instance j of the scene's TLS as num_surfaces + j
 * (-1 on a miss).  The surface part of intersectAll (renderer/common.h:
 * 1606-1620), exposed for tests and tools. */
VRT_API int vrt_intersect_surfaces(vrt_context ctx, const vrt_scene *scene, int n,
                                   const float *rays, int any, float *t_out, int *prim_out,
                                   int *surf_out);

VRT_API int vrt_delta_track(vrt_context ctx, const vrt_scene *scene, int n,
                            const float *rays, uint32_t *rng, float *t_out);

/* ---------------------------------------------------------------------- */
/* Renderers                                                               */
/* ---------------------------------------------------------------------- */

/* relMSE of a frame's float4 running mean against an FP64 reference given
 * as per-pixel sums (double4) times ref_scale: the mean over pixels and RGB
 * of (x - r)^2 / (r^2 + eps).  Synchronises; result in *out. */
VRT_API int vrt_relmse(vrt_context ctx, const float *accum, const double *ref_sum,
                       double ref_scale, int num_pixels, float eps, double *out);

/* Clear accum (and accum64, depth) of a frame: the accumulation reset of
 * Frame::renderFrame (frame/Frame.cpp:287-312). */
VRT_API int vrt_frame_reset(vrt_context ctx, const vrt_frame *frame);

/* 'raycast' renderer, one frame (renderer/Raycast_impl.cpp:245-322). */
VRT_API int vrt_raycast(vrt_context ctx, const vrt_scene *scene,
                        const vrt_camera *camera, const vrt_params *params,
                        const vrt_frame *frame, int frame_id);

/* 'pathtrace' renderer, one frame at accumulation index accum_id
 * (renderer/Pathtrace_impl.cpp:489-641). */
VRT_API int vrt_pathtrace(vrt_context ctx, const vrt_scene *scene,
                          const vrt_camera *camera, const vrt_params *params,
                          const vrt_frame *frame, int accum_id, int frame_id);

/* 'pathtrace' renderer, one frame, allocating nothing: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops).  Same arguments, semantics and image as
 * vrt_pathtrace; the persistent kernel's work-queue counters (16 uint32) live
 * in a file-scope device array of the library instead of the context's
 * allocation, cleared by a kernel at the start of every call (no memset).
 * Each GPU holds one such array until the process ends, so calls that
 * overlap on one GPU (several contexts or streams) race; serialise them.
 * Images are limited to VRT_NOALLOC_MAX_PIXELS pixels (the counters' range);
 * larger frames return VRT_ERROR_INVALID_ARGUMENT. */
#define VRT_NOALLOC_MAX_PIXELS (1LL << 30)
VRT_API int vrt_pathtrace_noalloc(vrt_context ctx, const vrt_scene *scene,
                                  const vrt_camera *camera, const vrt_params *params,
                                  const vrt_frame *frame, int accum_id, int frame_id);

/* ---------------------------------------------------------------------- */
/* Neural radiance cache (no counterpart in the reference)                 */
/* ---------------------------------------------------------------------- */
/* A small MLP mapping a scattering vertex (position in [0,1]^3, direction
 * as (theta/pi, phi/2pi), local density) to outgoing RGB radiance.
 * Encoding: each of the 5 position/direction scalars x becomes
 * sin(2^k pi x), cos(2^k pi x), k = 0..11 (120 features), plus the density
 * (121 inputs).  Network: 5 hidden layers of 64 ReLU units, 3 linear
 * outputs; fp16 or bf16 weights and activations on the matrix cores with
 * fp32 accumulation; fp32 master weights trained with Adam on the relative
 * L2 loss sum (p-t)^2 / (p^2 + 0.01) (p^2 not differentiated).
 * Parameters are one fp32 vector of vrt_nrc_param_count() values: padded
 * weights [layer][out][in] (layer 0: 64x128, layers 1-4: 64x64, layer 5:
 * 16x64) followed by padded biases (5 x 64 + 16). */
#define VRT_NRC_FP16 0
#define VRT_NRC_BF16 1
#define VRT_NRC_RAW_INPUTS 6

typedef struct vrt_nrc_config {
  int precision;        /* VRT_NRC_FP16 | VRT_NRC_BF16 */
  float learning_rate;  /* 1e-2 */
  float beta1, beta2;   /* 0.9, 0.99 */
  float epsilon;        /* 1e-8 */
  int max_batch;        /* largest n of one training step */
  uint64_t seed;        /* weight initialisation */
} vrt_nrc_config;

typedef struct vrt_nrc_s *vrt_nrc;

VRT_API int vrt_nrc_param_count(void);
VRT_API int vrt_nrc_create(vrt_context ctx, const vrt_nrc_config *cfg, vrt_nrc *out);
VRT_API int vrt_nrc_destroy(vrt_nrc nrc);
/* host fp32 parameter vector in / out (also re-emits the 16-bit copies) */
VRT_API int vrt_nrc_get_params(vrt_nrc nrc, float *params);
VRT_API int vrt_nrc_set_params(vrt_nrc nrc, const float *params);
/* n raw inputs (device, n x 6) -> n RGB predictions (device, n x 3) */
VRT_API int vrt_nrc_infer(vrt_nrc nrc, int n, const float *raw, float *rgb);
/* gradient sums over n samples (device inputs/targets n x 3); optional
 * device output of vrt_nrc_param_count() floats; loss sum to *loss */
VRT_API int vrt_nrc_gradients(vrt_nrc nrc, int n, const float *raw, const float *target,
                              float *grad, float *loss);
/* one Adam step on the mean gradient of n samples; loss sum to *loss */
VRT_API int vrt_nrc_train_step(vrt_nrc nrc, int n, const float *raw, const float *target,
                               float *loss);
/* 'pathtrace' frame with the radiance cache.  Paths end at their second
 * volume vertex (or their first: query_vertex 1) in a cache query (the cache
 * supplies the radiance leaving that vertex towards the path); a fraction of
 * paths, adapted on the GPU to give `train_samples` per frame, are traced on
 * instead (to full length, or train_extra_vertices further bounces and a
 * cache query of their own) and leave one training sample (the query
 * vertex's inputs, suffix radiance / throughput).
 * After the frame is resolved the cache takes `train_steps` Adam steps on
 * those samples.  Requires params->pixel_samples == 1 and a pinhole camera
 * (aperture_radius 0; VRT_ERROR_UNSUPPORTED otherwise).  The cache inputs
 * map world positions in [bounds_lo, bounds_hi] onto [0,1]^3. */
typedef struct vrt_nrc_render {
  int train;          /* 1: record training paths and train after the frame */
  int train_samples;  /* training samples per frame (65536) */
  int train_steps;    /* Adam steps (mini-batches) per frame */
  float bounds_lo[3];
  float bounds_hi[3];
  int query_vertex;   /* volume vertex that queries the cache: 2 (0 = default) or 1;
                         1 writes no g-buffer channels (rejected when the frame has them) */
  /* self-training (Mueller et al. 2021): 0 (default) traces training paths
     to full length; k in [1, 64] ends a training path whose vertex k bounces
     after its query vertex is a volume vertex in a cache query there, whose
     prediction (times the throughput gathered since the query vertex)
     completes both the pixel and the training target; paths meeting a
     surface or a light there go on to full length */
  int train_extra_vertices;
} vrt_nrc_render;

VRT_API int vrt_pathtrace_nrc(vrt_context ctx, const vrt_scene *scene,
                              const vrt_camera *camera, const vrt_params *params,
                              const vrt_frame *frame, vrt_nrc nrc,
                              const vrt_nrc_render *opts, int accum_id, int frame_id);
/* counts of the last cache frame: [0] queries, [1] training-slot claims,
 * [2] paths reaching vertex 2, [3] the previous frame's [2] (synchronises) */
VRT_API int vrt_nrc_frame_counts(vrt_nrc nrc, uint32_t counts[4]);

/* matrix-core lane-layout check of the 16x16x32 instruction the cache
 * kernels use: c (16x16) = a (16x32) b (32x16), device row-major fp32
 * in/out, operands rounded to fp16 or bf16 */
VRT_API int vrt_mfma_probe(vrt_context ctx, const float *a, const float *b, float *c,
                           int precision);

#ifdef __cplusplus
}
#endif
#endif
