"""The radiance cache inside the path tracer.

  zero cache         == plain path tracer with max_bounce 0 (bit-exact): the
                        cache replaces exactly the radiance from vertex 2 on
  constant cache c   adds throughput(vertex 2) * c = albedo * c per query
  all-training frame == plain full-length path tracer (bit-exact): training
                        paths are the plain paths (queries at vertex 2 or 1)
  first-vertex query  a constant cache c adds exactly c per query (throughput
                        1 before the first scattering)
  self-training      an all-training frame whose paths end in their own query
                        k bounces later equals, with a zero cache, the plain
                        path tracer cut at max_bounce k (bit-exact)
  learning gain      after online training the cached estimator is far
                        closer to the converged image than the truncated
                        estimator without the cache

The zero-cache, constant-cache and learning-gain checks run on a structured
fBm field with the fp16 cache and on a tetrahedral field (the unstructured-
plus-cache configuration) with the bf16 cache.
"""
import numpy as np
import pytest

from test_gpu_render import relmse

pytestmark = pytest.mark.gpu

NW = 8192 + 4 * 4096 + 1024


def _wl(pkg, n=32, w=96, h=64, max_bounce=8, field="fbm"):
    if field == "tets":
        # jittered lattice of 6 n^3 tetrahedra, the 2M-tet scene's generator
        return pkg.workloads.tets_pathtrace(n=max(4, n // 3), width=w, height=h, frames=1,
                                            max_bounce=max_bounce, grid=16)
    return pkg.workloads.fbm_pathtrace(n=n, width=w, height=h, frames=1, brick=8,
                                       max_bounce=max_bounce)


FIELDS = [("fbm", "fp16"), ("tets", "bf16")]


def _nrc(pkg, ctx, precision, **kw):
    vrt = pkg.vrt
    return vrt.Nrc(ctx, precision={"fp16": vrt.NRC_FP16, "bf16": vrt.NRC_BF16}[precision], **kw)


def _frame(gs, render, accum_id):
    gs.frame.reset()
    render(accum_id)
    gs.ctx.sync()
    return gs.frame.color_array()


@pytest.mark.parametrize("field,precision", FIELDS)
def test_zero_cache_equals_vertex1_truncation(pkg, gpu_ctx, field, precision):
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = _nrc(pkg, gpu_ctx, precision)
    try:
        nrc.set_params(np.zeros(nrc.nparams, np.float32))
        gs = wls.GpuScene(gpu_ctx, _wl(pkg, field=field))
        ref = wls.GpuScene(gpu_ctx, _wl(pkg, max_bounce=0, field=field))
        opts = gs.nrc_options(train=False)
        for a in range(3):
            g = _frame(gs, lambda k: gs.render_nrc(nrc, opts, k), a)
            r = _frame(ref, ref.render, a)
            np.testing.assert_array_equal(g, r)
        q, t, v2, _ = nrc.frame_counts()
        assert q == v2 > 100 and t == 0
    finally:
        nrc.close()


@pytest.mark.parametrize("field,precision", FIELDS)
def test_constant_cache_adds_throughput_times_value(pkg, gpu_ctx, field, precision):
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = _nrc(pkg, gpu_ctx, precision)
    try:
        c = np.array([0.25, 0.5, 1.0], np.float32)
        p = np.zeros(nrc.nparams, np.float32)
        p[NW + 5 * 64:NW + 5 * 64 + 3] = c  # output biases (exact in fp32: biases stay fp32)
        nrc.set_params(p)
        gs = wls.GpuScene(gpu_ctx, _wl(pkg, field=field))
        ref = wls.GpuScene(gpu_ctx, _wl(pkg, max_bounce=0, field=field))
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


@pytest.mark.parametrize("query_vertex", [2, 1])
def test_all_training_frame_equals_plain(pkg, gpu_ctx, query_vertex):
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = vrt.Nrc(gpu_ctx, max_batch=65536)
    try:
        gs = wls.GpuScene(gpu_ctx, _wl(pkg))
        ref = wls.GpuScene(gpu_ctx, _wl(pkg))
        # first frame: no previous count, every path reaching the query
        # vertex trains
        opts = gs.nrc_options(train=True, train_samples=65536, train_steps=1,
                              query_vertex=query_vertex)
        g = _frame(gs, lambda k: gs.render_nrc(nrc, opts, k), 0)
        r = _frame(ref, ref.render, 0)
        np.testing.assert_array_equal(g, r)
        q, t, v2, _ = nrc.frame_counts()
        assert q == 0 and t == v2 > 100
    finally:
        nrc.close()


def test_first_vertex_constant_cache(pkg, gpu_ctx):
    """Queries at the first volume vertex: the path has not scattered yet
    (throughput 1) and gathered nothing, so a constant cache c adds exactly c
    to every queried pixel over the zero cache's frame, which holds only the
    background of the pixels that miss the volume."""
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = vrt.Nrc(gpu_ctx)
    try:
        gs = wls.GpuScene(gpu_ctx, _wl(pkg))
        opts = gs.nrc_options(train=False, query_vertex=1)
        nrc.set_params(np.zeros(nrc.nparams, np.float32))
        z = _frame(gs, lambda k: gs.render_nrc(nrc, opts, k), 0)
        c = np.array([0.25, 0.5, 1.0], np.float32)
        p = np.zeros(nrc.nparams, np.float32)
        p[NW + 5 * 64:NW + 5 * 64 + 3] = c  # output biases
        nrc.set_params(p)
        g = _frame(gs, lambda k: gs.render_nrc(nrc, opts, k), 0)
        d = (g - z)[..., :3]
        queried = np.abs(d[..., 2]) > 1e-6
        np.testing.assert_allclose(d[queried], np.broadcast_to(c, d[queried].shape), rtol=1e-6)
        assert np.all(d[~queried] == 0)
        q, t, v1, _ = nrc.frame_counts()
        assert queried.sum() == q == v1 > 100 and t == 0
        assert np.all(z[queried][..., :3] == 0)
    finally:
        nrc.close()


def test_first_vertex_queries_reject_gbuffer(pkg, gpu_ctx):
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = vrt.Nrc(gpu_ctx)
    try:
        gs = wls.GpuScene(gpu_ctx, _wl(pkg), gbuffer=True)
        with pytest.raises(vrt.VrtError, match="g-buffer"):
            gs.render_nrc(nrc, gs.nrc_options(train=False, query_vertex=1), 0)
        gs.render_nrc(nrc, gs.nrc_options(train=False, query_vertex=2), 0)
        with pytest.raises(vrt.VrtError, match="query_vertex"):
            gs.render_nrc(nrc, gs.nrc_options(train=False, query_vertex=3), 0)
    finally:
        nrc.close()


@pytest.mark.parametrize("textra", [1, 2])
def test_self_training_zero_cache_equals_truncation(pkg, gpu_ctx, textra):
    """First frame, every path reaching vertex 2 trains; with self-training a
    training path whose vertex textra bounces later is a volume vertex ends
    there in a cache query.  A zero cache adds nothing, so the frame is the
    plain path tracer's with max_bounce textra (the zero cache at vertex 2
    itself gives max_bounce 0).  A constant cache c then adds throughput * c
    at each terminal query: the same multiple of c in every channel (the
    medium is grey)."""
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = vrt.Nrc(gpu_ctx, max_batch=65536)
    try:
        nrc.set_params(np.zeros(nrc.nparams, np.float32))
        gs = wls.GpuScene(gpu_ctx, _wl(pkg))
        ref = wls.GpuScene(gpu_ctx, _wl(pkg, max_bounce=textra))
        opts = gs.nrc_options(train=True, train_samples=65536, train_steps=1,
                              train_extra_vertices=textra)
        g = _frame(gs, lambda k: gs.render_nrc(nrc, opts, k), 0)
        r = _frame(ref, ref.render, 0)
        np.testing.assert_array_equal(g, r)
        q, t, v2, _ = nrc.frame_counts()
        assert t == v2 > 100 and 0 < q < t

        c = np.array([0.25, 0.5, 1.0], np.float32)
        p = np.zeros(nrc.nparams, np.float32)
        p[NW + 5 * 64:NW + 5 * 64 + 3] = c  # output biases
        nrc2 = vrt.Nrc(gpu_ctx, max_batch=65536)  # fresh counters: all paths train again
        try:
            nrc2.set_params(p)
            g2 = _frame(gs, lambda k: gs.render_nrc(nrc2, opts, k), 0)
            q2 = nrc2.frame_counts()[0]
        finally:
            nrc2.close()
        d = (g2 - r)[..., :3]
        queried = np.abs(d[..., 2]) > 1e-6
        assert queried.sum() == q2 == q
        k = d[queried] / c
        np.testing.assert_allclose(k[:, 0], k[:, 2], rtol=1e-5)
        np.testing.assert_allclose(k[:, 1], k[:, 2], rtol=1e-5)
        assert np.all(np.abs(d[~queried]) <= 1e-6)
    finally:
        nrc.close()


@pytest.mark.parametrize("textra", [0, 64])
def test_self_training_option_bounds(pkg, gpu_ctx, textra):
    vrt, wls = pkg.vrt, pkg.workloads
    nrc = vrt.Nrc(gpu_ctx)
    try:
        gs = wls.GpuScene(gpu_ctx, _wl(pkg))
        gs.render_nrc(nrc, gs.nrc_options(train_samples=1024, train_extra_vertices=textra), 0)
        for bad in (-1, 65):
            with pytest.raises(vrt.VrtError, match="train_extra_vertices"):
                gs.render_nrc(nrc, gs.nrc_options(train_extra_vertices=bad), 0)
    finally:
        nrc.close()


@pytest.mark.parametrize("field,precision,query_vertex,textra",
                         [("fbm", "fp16", 2, 0), ("fbm", "fp16", 1, 0), ("tets", "bf16", 2, 0),
                          ("fbm", "fp16", 2, 2), ("tets", "bf16", 2, 1)])
def test_online_training_gain(pkg, gpu_ctx, field, precision, query_vertex, textra):
    """relMSE against the converged plain image: the trained cache beats the
    truncated estimator without a cache (its paths end after the first
    vertex's light sample) by a wide margin, whether it is queried at the
    second volume vertex or at the first, where it also stands in for that
    light sample, and whether its training paths run to full length or end
    in a query of their own (self-training, textra > 0)."""
    vrt, wls = pkg.vrt, pkg.workloads
    wl = _wl(pkg, n=48, w=128, h=96, field=field)
    ref = wls.GpuScene(gpu_ctx, wl, accum64=True)
    ref.frame.reset()
    nref = 2048
    for a in range(nref):
        ref.render(a)
    gpu_ctx.sync()
    truth = ref.frame.accum64_array() / nref

    trunc = wls.GpuScene(gpu_ctx, _wl(pkg, n=48, w=128, h=96, max_bounce=0, field=field),
                         volume=ref.volume if field == "tets" else None)
    trunc.frame.reset()
    for a in range(256):
        trunc.render(a)
    gpu_ctx.sync()
    e_trunc = relmse(trunc.frame.accum_array(), truth)

    nrc = _nrc(pkg, gpu_ctx, precision, max_batch=4096, seed=7)
    try:
        gs = wls.GpuScene(gpu_ctx, wl, volume=ref.volume)
        opts = gs.nrc_options(train=True, train_samples=4096, train_steps=2,
                              query_vertex=query_vertex, train_extra_vertices=textra)
        for a in range(300):  # warm-up training, frames discarded
            gs.render_nrc(nrc, opts, 10000 + a)
        gs.frame.reset()
        for a in range(256):
            gs.render_nrc(nrc, opts, a)
        gpu_ctx.sync()
        e_nrc = relmse(gs.frame.accum_array(), truth)
    finally:
        nrc.close()
    print(f"{field} {precision}: relMSE truncated {e_trunc:.4f}  with trained cache "
          f"{e_nrc:.4f} (queries at vertex {query_vertex}, self-training {textra})")
    assert e_nrc < 0.3 * e_trunc
