"""The radiance cache inside the path tracer.

  zero cache         == plain path tracer with max_bounce 0 (bit-exact): the
                        cache replaces exactly the radiance from vertex 2 on
  constant cache c   adds throughput(vertex 2) * c = albedo * c per query
  all-training frame == plain full-length path tracer (bit-exact): training
                        paths are the plain paths
  learning gain      after online training the cached estimator is far
                        closer to the converged image than the truncated
                        estimator without the cache
"""
import numpy as np
import pytest

from test_gpu_render import relmse

pytestmark = pytest.mark.gpu

NW = 8192 + 4 * 4096 + 1024


def _wl(pkg, n=32, w=96, h=64, max_bounce=8):
    return pkg.workloads.fbm_pathtrace(n=n, width=w, height=h, frames=1, brick=8,
                                       max_bounce=max_bounce)


def _frame(gs, render, accum_id):
    gs.frame.reset()
    render(accum_id)
    gs.ctx.sync()
    return gs.frame.color_array()


def test_zero_cache_equals_vertex1_truncation(pkg, gpu_ctx):
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = vrt.Nrc(gpu_ctx)
    try:
        nrc.set_params(np.zeros(nrc.nparams, np.float32))
        gs = wls.GpuScene(gpu_ctx, _wl(pkg))
        ref = wls.GpuScene(gpu_ctx, _wl(pkg, max_bounce=0))
        opts = gs.nrc_options(train=False)
        for a in range(3):
            g = _frame(gs, lambda k: gs.render_nrc(nrc, opts, k), a)
            r = _frame(ref, ref.render, a)
            np.testing.assert_array_equal(g, r)
        q, t, v2, _ = nrc.frame_counts()
        assert q == v2 > 100 and t == 0
    finally:
        nrc.close()


def test_constant_cache_adds_throughput_times_value(pkg, gpu_ctx):
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = vrt.Nrc(gpu_ctx)
    try:
        c = np.array([0.25, 0.5, 1.0], np.float32)
        p = np.zeros(nrc.nparams, np.float32)
        p[NW + 5 * 64:NW + 5 * 64 + 3] = c  # output biases
        nrc.set_params(p)
        gs = wls.GpuScene(gpu_ctx, _wl(pkg))
        ref = wls.GpuScene(gpu_ctx, _wl(pkg, max_bounce=0))
        opts = gs.nrc_options(train=False)
        # accum_id 0 after a reset: the frame holds the sample itself
        g = _frame(gs, lambda k: gs.render_nrc(nrc, opts, k), 0)
        r = _frame(ref, ref.render, 0)
        d = (g - r)[..., :3]
        queried = np.abs(d[..., 2]) > 1e-6
        # throughput at vertex 2 is the albedo (0.8) of vertex 1
        np.testing.assert_allclose(d[queried], np.broadcast_to(0.8 * c, d[queried].shape),
                                   rtol=1e-5)
        assert np.all(np.abs(d[~queried]) <= 1e-6)
        assert queried.sum() == nrc.frame_counts()[0]
    finally:
        nrc.close()


def test_all_training_frame_equals_plain(pkg, gpu_ctx):
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = vrt.Nrc(gpu_ctx, max_batch=65536)
    try:
        gs = wls.GpuScene(gpu_ctx, _wl(pkg))
        ref = wls.GpuScene(gpu_ctx, _wl(pkg))
        # first frame: no previous count, every vertex-2 path trains
        opts = gs.nrc_options(train=True, train_samples=65536, train_steps=1)
        g = _frame(gs, lambda k: gs.render_nrc(nrc, opts, k), 0)
        r = _frame(ref, ref.render, 0)
        np.testing.assert_array_equal(g, r)
        q, t, v2, _ = nrc.frame_counts()
        assert q == 0 and t == v2 > 100
    finally:
        nrc.close()


def test_online_training_gain(pkg, gpu_ctx):
    """relMSE against the converged plain image: the trained cache beats the
    same truncated estimator without a cache by a wide margin."""
    vrt, wls = pkg.vrt, pkg.workloads
    wl = _wl(pkg, n=48, w=128, h=96)
    ref = wls.GpuScene(gpu_ctx, wl, accum64=True)
    ref.frame.reset()
    nref = 2048
    for a in range(nref):
        ref.render(a)
    gpu_ctx.sync()
    truth = ref.frame.accum64_array() / nref

    trunc = wls.GpuScene(gpu_ctx, _wl(pkg, n=48, w=128, h=96, max_bounce=0))
    trunc.frame.reset()
    for a in range(256):
        trunc.render(a)
    gpu_ctx.sync()
    e_trunc = relmse(trunc.frame.accum_array(), truth)

    nrc = vrt.Nrc(gpu_ctx, max_batch=4096, seed=7)
    try:
        gs = wls.GpuScene(gpu_ctx, wl)
        opts = gs.nrc_options(train=True, train_samples=4096, train_steps=2)
        for a in range(300):  # warm-up training, frames discarded
            gs.render_nrc(nrc, opts, 10000 + a)
        gs.frame.reset()
        for a in range(256):
            gs.render_nrc(nrc, opts, a)
        gpu_ctx.sync()
        e_nrc = relmse(gs.frame.accum_array(), truth)
    finally:
        nrc.close()
    print(f"relMSE truncated {e_trunc:.4f}  with trained cache {e_nrc:.4f}")
    assert e_nrc < 0.3 * e_trunc
