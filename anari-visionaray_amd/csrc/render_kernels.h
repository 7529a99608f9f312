// Launchers of the gfx950 rendering kernels (render_kernels.hip).
#pragma once

#include "kernel_common.h"

namespace vrt {

hipError_t launch_raycast(const KScene &sc, const KCamera &cam, const KParams &P, const KFrame &F,
                          uint32_t *queues, uint32_t *queues_next, int num_cus, hipStream_t s);
// persistent path tracer: `blocks` workgroups (a multiple of 8) drain the
// per-XCD pixel queues; `queues` must be zero, `queues_next` is zeroed for
// the following launch
// the following launch.  With `nrc`, paths end in radiance-cache queries at
// their second volume vertex (see KNrc) and the frame is written by
// launch_nrc_resolve once the queries are answered.
hipError_t launch_pathtrace(const KScene &sc, const KCamera &cam, const KParams &P,
                            const KFrame &F, int accumID, uint32_t *queues,
                            uint32_t *queues_next, int blocks, hipStream_t s,
                            const KNrc *nrc = nullptr);
// the device address of the file-scope work-queue counters on the current
// device (vrt_pathtrace_noalloc), and the path tracer launched over them:
// the counters are cleared by a kernel first, nothing is allocated
hipError_t static_queue_address(uint32_t **ptr);
hipError_t launch_pathtrace_static(const KScene &sc, const KCamera &cam, const KParams &P,
                                   const KFrame &F, int accumID, uint32_t *scratch_ptr,
                                   int blocks, hipStream_t s);
// self-training: add the terminal queries' predictions to the targets
hipError_t launch_nrc_train_targets(const KNrc &N, hipStream_t s);
hipError_t launch_nrc_resolve(const KFrame &F, const KNrc &N, int accumID, hipStream_t s);
hipError_t launch_taa_resolve(const KFrame &F, float alpha, hipStream_t s);
// adds sum (x - r)^2 / (r^2 + eps) over n pixels' RGB to *out
hipError_t launch_relmse(const float4 *x, const double *ref_sum, double ref_scale, int n,
                         float eps, double *out, hipStream_t s);
int pathtrace_blocks_per_cu(const KScene &sc, size_t lds_bytes, bool nrc = false,
                            bool grad = false);
hipError_t launch_bricks(const float *lin, const int dims[3], float *bricks, hipStream_t s);
hipError_t launch_value_ranges(const KVolume &v, float *ranges, hipStream_t s);
// nanovdb: min/max (float2) of each block slot's values
hipError_t launch_nanovdb_slot_ranges(const int32_t *slots, size_t nslots, const float *leaves,
                                      const float *consts, float *slot_ranges, hipStream_t s);
// nanovdb: cell value ranges from the cells' voxel boxes (6 ints per cell:
// lo xyz, hi xyz inclusive) over the slot ranges
hipError_t launch_nanovdb_value_ranges(const KVolume &v, const int32_t *cell_voxels,
                                       float *ranges, hipStream_t s);
// unstructured: cell value ranges rasterized from the tetrahedra's bounding
// boxes and value ranges (cells without one keep the empty range)
hipError_t launch_tet_value_ranges(const KVolume &v, float *ranges, hipStream_t s);
hipError_t launch_majorants(const KVolume &v, const float *ranges, float *maj, uint8_t *bytes,
                            float *step, hipStream_t s);
// LDS bytes a render kernel needs for the scene's tables and majorant grids
size_t scene_lds_size(const KScene &sc);
// LDS bytes of the path tracer: the above plus the BVH traversal stacks and
// the staged BVH nodes when the scene has surfaces or quad lights
size_t pathtrace_lds_size(const KScene &sc);
// the path tracer's pass schedule for the scene (vrt_params.pass_schedule
// 1-3), and its KParams fields
int pathtrace_schedule(const KScene &sc);
void set_pass_schedule(KParams &P, int schedule);
// sets the BVH nodes each workgroup stages in LDS (sc.staged[i].lds_nodes):
// as many top nodes as fit beside `base` bytes per workgroup without
// lowering the `per_cu` resident workgroups
void plan_staged_nodes(KScene &sc, size_t base, int per_cu);
hipError_t launch_frame_reset(const KFrame &F, hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
hipError_t launch_sample_field(const KVolume &v, int n, const float *pts, float *out,
                               hipStream_t s);
hipError_t launch_intersect_surfaces(const KScene &sc, int n, const float *rays, int any,
                                     float *t_out, int *prim_out, int *surf_out, hipStream_t s);
hipError_t launch_delta_track(const KScene &sc, int n, const float *rays, uint32_t *rng,
                              float *tout, hipStream_t s);

}  // namespace vrt
