/*
 * ORACLE -- test infrastructure, not product code.
 *
 * CPU restatement of anari-visionaray's per-pixel volume rendering path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / the timed CPU baseline.
 *
 * Every routine cites the reference file:line it restates (paths relative to
 * the reference repository root).  Functions marked "parity unpinned" restate
 * behaviour that lives in the (absent) visionaray library; they follow the
 * ANARI specification / textbook definition named next to them.
 *
 * Precision: every renderer is instantiated twice, FP32 ("parity mode", the
 * arithmetic the reference performs) and FP64 ("oracle mode", the converged
 * reference used for relMSE figures).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LIGHT_DIRECTIONAL 0
#define ORC_LIGHT_QUAD 1
#define ORC_LIGHT_HDRI 2

typedef struct orc_light {
  int type;      /* ORC_LIGHT_* */
  int visible;
  float dir[3];  /* directional: ANARI "direction" (direction light travels) */
  float pos[3];  /* quad: "position" */
  float e1[3];   /* quad: "edge1" */
  float e2[3];   /* quad: "edge2" */
  float color[3];
  float intensity; /* directional: "irradiance", quad: "intensity" */
  /* hdri (HDRI.cpp:102-162; one per scene): RGB float map, x fastest,
     "up", "scale"; `dir` holds its "direction" */
  const float *hdri_rgb;
  int hdri_w, hdri_h;
  float up[3];
  float scale;
} orc_light;

typedef struct orc_volume {
  /* structuredRegular field, x-fastest linear float voxels */
  const float *voxels;
  int dims[3];
  float origin[3];
  float spacing[3];
  /* transferFunction1D: resampled float4 table (see orc_tf_build) */
  const float *tf_rgba;
  int tf_size;
  float value_lo, value_hi;
  float unit_distance;
  /* majorant grid over voxel space [0,1]^3; NULL -> one segment, majorant 1
     (the reference GPU build's behaviour, DeviceCopyableObjects.h:724,777) */
  const float *majorants;
  int grid_dims[3];
  /* instance transform object->world, row-major 3x4 [R|t] */
  float xfm[12];
  /* field type: 0 structuredRegular (above), 1 nanovdb: `nanovdb` is a
     NanoVDB float grid buffer (NanoVDBField.cpp), filter 0 linear,
     1 nearest; bounds, voxel space and cell size then come from the grid */
  int field_type;
  const void *nanovdb;
  int filter;
  /* field type 2, unstructured: the ANARI arrays of UnstructuredField
     (UnstructuredField.cpp:17-27): float3 positions, per-vertex values
     (NULL: none), indices, first index per cell, VTK/barney cell types,
     per-cell values (NULL: none).  Only tetrahedral cells are restated;
     bounds, voxel space (identity) and cell size come from the cells. */
  const float *vertex_position;
  const float *vertex_data;
  int64_t num_vertices;
  const uint32_t *index;
  int64_t num_indices;
  const uint32_t *cell_index;
  const uint8_t *cell_type;
  const float *cell_data;
  int64_t num_cells;
  /* ANARI "id" of the volume and of its instance (Volume.cpp:25, the
     instance's userID), reported in the g-buffer; ~0u when unset */
  uint32_t object_id;
  uint32_t instance_id;
} orc_volume;

typedef struct orc_trimesh {
  const float *vertices; /* float3 per vertex */
  const uint32_t *indices; /* uint3 per triangle; NULL -> soup */
  int num_triangles;
  float color[3]; /* matte color */
  float opacity;
  uint32_t object_id;   /* Surface "id" (Surface.cpp:21), ~0u when unset */
  uint32_t instance_id; /* the instance's "id" */
} orc_trimesh;

/* A triangle mesh placed by an instance (dco::Instance,
   DeviceCopyableObjects.h:1659-1680; intersect(Ray, Instance), :1772-1814):
   rays meet it in object space, ori' = affine_inv (ori + trans_inv),
   dir' = affine_inv dir, and its normals go back by normal_xfm =
   transpose(affine_inv) (Instance.cpp:94-97, Pathtrace_impl.cpp:243-247).
   The matrices (row-major) are given, so the hits depend on traversal and
   arithmetic alone.  A scene's surface hit is the closest over its meshes
   (identity) and its instances: a mesh keeps ties with an instance, and
   among instances the lower index wins. */
/* affine_inv = inverse(L) of a row-major 3x4 [L | t] (cofactors in
   double, rounded to float), trans_inv = -t, normal_xfm =
   transpose(affine_inv) (Instance.cpp:94-97); returns 0, or -1 for a
   singular or non-finite L */
int orc_instance_transforms(const float xfm[12], float affine_inv[9], float trans_inv[3],
                            float normal_xfm[9]);

typedef struct orc_instance {
  orc_trimesh mesh;
  float affine_inv[9];
  float trans_inv[3];
  float normal_xfm[9];
} orc_instance;

typedef struct orc_scene {
  int num_volumes;
  const orc_volume *volumes;
  int num_lights;
  const orc_light *lights;
  int num_meshes;
  const orc_trimesh *meshes;
  int num_instances;
  const orc_instance *instances;
} orc_scene;

typedef struct orc_camera {
  float pos[3], dir[3], up[3];
  float fovy;   /* radians */
  float aspect;
  float region[4]; /* imageRegion: lo.x lo.y hi.x hi.y */
  float shutter[2];
  float aperture_radius; /* thin lens; 0: pinhole */
  float focus_distance;  /* plane in focus, along the view direction */
} orc_camera;

typedef struct orc_params {
  float bg[4];
  float ambient_color[3];
  float ambient_radiance;
  float occlusion_distance;
  int ambient_samples;
  int max_bounce;
  int pixel_samples;
  float volume_sampling_rate; /* raycast only */
  /* pixel streams seeded with (pixelID, accumID + rng_offset); 0 is the
     reference's seeding */
  uint32_t rng_offset;
  /* clip planes (N, D): each ray keeps N.x <= D (renderer/common.h:1469-1482) */
  int num_clip_planes;
  float clip_planes[8][4];
  /* Henyey-Greenstein anisotropy of the volume phase function (an
     extension; 0 = the reference's isotropic phase) */
  float phase_g;
  /* gradientShading (Renderer.cpp:50): raycast matte shading along the
     central-difference gradient; path-tracer volume normal and matte NEE */
  int gradient_shading;
  /* extension: 1 = shadow and AO rays estimate the volumes' transmittance
     by ratio tracking instead of the reference's binary delta tracking */
  int ratio_tracking;
} orc_params;

/* --- A1: LCG<4> / Random (common.h:55-102) --- */
uint32_t orc_tea_seed(uint32_t v0, uint32_t v1);
void orc_lcg_stream(uint32_t seed0, uint32_t seed1, int n, float *out);

/* --- A13: TransferFunction1D table build (TransferFunction1D.cpp:66-111) --- */
int orc_tf_build(const float *colors, int num_colors, int color_channels,
                 const float *opacities, int num_opacities,
                 const float uniform_rgba[4], float *out_rgba);

/* --- A10/A11/A12: sampleField at voxel-space points (NaN where the field
   has no value) --- */
void orc_sample_field(const orc_volume *vol, int n, const float *p_voxel,
                      float *out);
void orc_post_classify(const orc_volume *vol, int n, const float *values,
                       float *out_rgba);

/* --- A12: unstructured field.  Info: bounds (lo xyz, hi xyz), cell size
   (the running avgCellDiagonal, UnstructuredField.cpp:120-170), cells kept /
   skipped; returns 0, or -1 if the field has non-tetrahedral cells.  The
   brute-force sampler tests every cell in id order (the pin of the oracle's
   own point-location grid); NaN where no cell contains the point. --- */
int orc_unstructured_info(const orc_volume *vol, float bounds[6], float *cell_size,
                          int64_t *kept, int64_t *skipped);
void orc_sample_field_brute(const orc_volume *vol, int n, const float *p, float *out);

/* --- A9: majorant grid (conservative), and the reference rasterizer --- */
void orc_build_majorants(const orc_volume *vol, int gx, int gy, int gz,
                         float *out);
void orc_build_majorants_reference(const orc_volume *vol, int gx, int gy,
                                   int gz, float *out);

/* --- A8: dda3 (renderer/DDA.h:14-112); returns #cells written --- */
int orc_dda_cells(const float ori[3], const float dir[3], float tmin,
                  float tmax, const int dims[3], const float lo[3],
                  const float hi[3], int max_cells, int *cells, float *t0,
                  float *t1);

/* --- A7: delta tracking along one world ray through one volume.
   Returns hit (0/1); t_out world distance; consumes rng state in place. --- */
int orc_delta_track(const orc_volume *vol, const float ori[3],
                    const float dir[3], float tmin, float tmax,
                    uint32_t *rng_state, float *t_out, int fp64);
/* ratio tracking of one ray (an extension for shadow / AO rays): the
   transmittance estimate Tr in [0, 1] */
float orc_ratio_track(const orc_volume *v, const float ori[3], const float dir[3], float tmin,
                      float tmax, uint32_t *state, int fp64);

/* --- A14/A15: raycast renderer (Raycast_impl.cpp:245-322) --- */
void orc_raycast(const orc_scene *scene, const orc_camera *cam,
                 const orc_params *params, int width, int height,
                 float *out_rgba, int threads);

/* --- A16: one path-traced frame: per-pixel sample colour (the value
   Pathtrace_impl.cpp:639 passes to writeSample), over rows [y0,y1) and
   columns with x % stride == 0 && y % stride == 0 when stride > 1. --- */
void orc_pathtrace_frame(const orc_scene *scene, const orc_camera *cam,
                         const orc_params *params, int width, int height,
                         int accum_id, int frame_id, int fp64, int stride,
                         int y0, int y1, float *out_rgba, int threads,
                         uint64_t *lookups);

/* --- A20: Frame::accumSample running mean over frames [0,frames) --- */
void orc_pathtrace_accumulate(const orc_scene *scene, const orc_camera *cam,
                              const orc_params *params, int width, int height,
                              int first_accum_id, int frames, int fp64,
                              int stride, double *out_rgba, int threads);

/* cvt_uint32 / cvt_uint32_srgb (frame/common.h:31-47) */
uint32_t orc_cvt_uint32(const float rgba[4]);
uint32_t orc_cvt_uint32_srgb(const float rgba[4]);

/* powf of the opacity correction from IEEE basic operations, and its
   64-entry log2 table {1/c_i, -log2(1/c_i)} */
float orc_pow_ieee(float x, float y);
void orc_pow_log2_table(float *out);

/* camera ray for pixel coordinate (x, y) (the pinhole of the ANARI perspective
   camera; pinned against the look-at / perspective projection in the tests) */
/* --- A5/A6: closest (any=0) or any (any=1) hit of n rays {ori, dir, tmin,
   tmax} against a mesh, through the oracle's BVH or by brute force; t = inf
   and prim = -1 on a miss --- */
void orc_mesh_intersect(const orc_trimesh *m, int n, const float *rays, int any, int brute,
                        float *t_out, int *prim_out);
/* Closest (any = 0) or any (any = 1) surface hit of n rays {ori xyz, dir
   xyz, tmin, tmax} in a scene's meshes and instances (float arithmetic):
   t (+inf on a miss), the triangle index in its mesh (-1) and what was hit
   (mesh i as i, instance j as num_meshes + j; -1). */
void orc_scene_intersect_surfaces(const orc_scene *sc, int n, const float *rays, int any,
                                  float *t_out, int *prim_out, int *surf_out);
/* --- A17: n quad-light samples seen from ref (RNG seeded (seed, i)):
   positions and pdfs --- */
void orc_quad_light_sample(const orc_light *L, const float ref[3], int n, uint32_t seed,
                           float *pos_out, float *pdf_out);
/* --- Henyey-Greenstein phase (extension): n directions sampled around d
   (RNG seeded (seed, i)), and 4 pi p_HG(cos theta) --- */
void orc_hg_sample(float g, const float d[3], int n, uint32_t seed, float *out);
double orc_hg_4pi(double g, double cos_theta);
/* --- isotropic phase sample (uniform_sample_sphere, Pathtrace_impl.cpp:335;
   parity unpinned: PBRT definition): n directions (RNG seeded (seed, i)) --- */
void orc_uniform_sample_sphere(int n, uint32_t seed, float *out);
/* --- Russian roulette of the path tracer (Pathtrace_impl.cpp:594-599): 1
   (default, the reference) or 0 (paths end only at max_bounce or a miss).
   Process-wide; a test switch for the estimator's unbiasedness. --- */
void orc_set_roulette(int enabled);
/* --- HDRI light (type ORC_LIGHT_HDRI): its CDFs (h row-total entries, w*h
   per-row entries; -1 for a degenerate light), n light samples (RNG seeded
   (seed, i)): direction, pdf and radiance; and for given directions the
   radiance, the hit-side light pdf and the map coordinates --- */
int orc_hdri_cdf(const orc_light *L, float *rows, float *cols);
void orc_hdri_sample(const orc_light *L, int n, uint32_t seed, float *dirs, float *pdfs,
                     float *rgb);
void orc_hdri_eval(const orc_light *L, int n, const float *dirs, float *rgb, float *pdfs,
                   float *uv);
void orc_camera_ray(const orc_camera *cam, float x, float y, int width,
                    int height, float *ori, float *dir);
/* the pinhole ray of (x, y) moved onto the thin lens by the lens sample
   (u1, u2), in FP64 */
void orc_camera_lens_ray(const orc_camera *cam, float x, float y, int width, int height,
                         float u1, float u2, double *ori, double *dir);

/* --- g-buffer of one frame (Frame::fillGBuffer, DeviceCopyableObjects.h:
   2745-2774), FP32.  Path tracer: per pixel the first radiance hit of the
   closest sample (Pathtrace_impl.cpp:226-285, 523-617); ray caster: the
   closest volume box's ids (Raycast_impl.cpp:18-20, 230-237).  `depth` is
   the caller's depth buffer (1e31 after an accumulation reset), read and
   updated: a pixel's channels are written where its depth <= the stored
   one.  normal / albedo float3, ids uint32 (~0u: none); any output but
   depth may be NULL. --- */
void orc_pathtrace_gbuffer(const orc_scene *scene, const orc_camera *cam,
                           const orc_params *params, int width, int height, int accum_id,
                           float *depth, float *normal, float *albedo, uint32_t *prim,
                           uint32_t *obj, uint32_t *inst, int threads);
/* --- one TAA frame (path tracer "taa", Frame.cpp:339-381), FP32: the
   frame's samples resolved against the history prev_rgba (w*h float4) /
   prev_albedo (w*h float3), which are then replaced by this frame's; the
   output colours (float4) and motion vectors (float2, may be NULL).
   prev_camera NULL: no motion. --- */
void orc_pathtrace_taa(const orc_scene *scene, const orc_camera *cam,
                       const orc_camera *prev_camera, const orc_params *params, int width,
                       int height, int accum_id, float alpha, float *prev_rgba, float *prev_albedo,
                       float *out_rgba, float *motion, int threads);
void orc_raycast_gbuffer(const orc_scene *scene, const orc_camera *cam,
                         const orc_params *params, int width, int height, float *depth,
                         float *normal, float *albedo, uint32_t *prim, uint32_t *obj,
                         uint32_t *inst, int threads);

#ifdef __cplusplus
}
#endif
