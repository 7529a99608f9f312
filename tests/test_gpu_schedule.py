"""The path tracer's pass schedule (vrt_params.pass_schedule) changes only
the order in which lanes trace their rays, never a pixel:

  chained (1) == one ray per pass (2) == one ray per pass with idle lanes
  refilled 16 at a time (3), bit for bit, over three accumulated
  frames on a structured fBm field, a tetrahedral field, a NanoVDB cloud and
  the mixed triangle-mesh + volume scene with a quad light; and for a
  radiance-cache frame (fixed weights, queries at the second vertex), whose
  queries and training samples the two schedules write in different orders.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes(pkg):
    wls = pkg.workloads
    return {
        "fbm": wls.fbm_pathtrace(n=32, width=96, height=64, frames=3, brick=8),
        "tets": wls.tets_pathtrace(n=10, width=96, height=64, frames=3, grid=16),
        "cloud": wls.cloud_nanovdb_pathtrace(n=128, width=96, height=64, frames=3),
        "mixed": wls.mixed_pathtrace(level=4, vol_n=32, width=96, height=64, frames=3),
    }


def _accum(gs, frames=3):
    gs.frame.reset()
    for a in range(frames):
        gs.render(a)
    gs.ctx.sync()
    return np.asarray(gs.frame.accum_array())


@pytest.mark.parametrize("scene", ["fbm", "tets", "cloud", "mixed"])
def test_schedules_render_the_same_image(pkg, gpu_ctx, scene):
    wl = _scenes(pkg)[scene]
    images = []
    for schedule in (1, 2, 3, 0):
        gs = pkg.workloads.GpuScene(gpu_ctx, wl, params={"pass_schedule": schedule})
        images.append(_accum(gs))
    for other in images[1:]:
        np.testing.assert_array_equal(images[0], other)
    assert float(images[0][..., :3].mean()) > 1e-3


def test_schedules_render_the_same_cache_frame(pkg, gpu_ctx):
    vrt, wls = pkg.vrt, pkg.workloads
    wl = wls.fbm_pathtrace(n=32, width=96, height=64, frames=1, brick=8)
    rng = np.random.default_rng(7)
    images = []
    for schedule in (1, 2, 3):
        nrc = vrt.Nrc(gpu_ctx, max_batch=16384)
        try:
            nrc.set_params((0.05 * rng.standard_normal(nrc.nparams)).astype(np.float32)
                           if not images else params)
            params = nrc.params()
            gs = wls.GpuScene(gpu_ctx, wl, params={"pass_schedule": schedule})
            opts = gs.nrc_options(train=False)
            gs.frame.reset()
            gs.render_nrc(nrc, opts, 0)
            gpu_ctx.sync()
            images.append(np.asarray(gs.frame.color_array()))
            assert nrc.frame_counts()[0] > 100
        finally:
            nrc.close()
    for other in images[1:]:
        np.testing.assert_array_equal(images[0], other)


def test_schedule_out_of_range_is_refused(pkg, gpu_ctx):
    wl = pkg.workloads.fbm_pathtrace(n=16, width=16, height=16, frames=1, brick=8)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, params={"pass_schedule": 4})
    with pytest.raises(pkg.vrt.VrtError, match="pass_schedule"):
        gs.render(0)
