"""GPU parity of the HIP rendering path against the CPU oracle.

Every test renders the same seeded inputs through the C ABI (libvrt.so) and
through the oracle's restatement of the reference, and compares:
  raycast   max-abs <= 1e-3 per pixel on [0,1] RGBA (deterministic renderer)
  pathtrace stream-aligned per-pixel agreement, and relMSE <= 1e-3 of the
            accumulated mean (eps = 1e-2 in the relMSE denominator)
  fetch     trilinear values within 2e-6 absolute; majorants bit-exact
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RELMSE_EPS = 1e-2


def relmse(x, r, eps=RELMSE_EPS):
    x = np.asarray(x, np.float64)[..., :3]
    r = np.asarray(r, np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def _gpu_volume(pkg, ctx, wl):
    return pkg.vrt.DeviceVolume(ctx, wl.voxels, wl.origin, wl.spacing, wl.tf, wl.value_range,
                                wl.unit_distance, wl.grid_dims)


def test_sample_field_ragged(pkg, oracle, gpu_ctx):
    vrt = pkg.vrt
    rng = np.random.default_rng(7)
    vox = rng.random((27, 33, 40), dtype=np.float32)  # (nz, ny, nx), not multiples of 8
    wl_tf = pkg.workloads.ramp_tf(16)
    dv = vrt.DeviceVolume(gpu_ctx, vox, (-0.5,) * 3, (0.025,) * 3, wl_tf, (0, 1), 0.1)
    ov = oracle.OracleVolume(vox, (-0.5,) * 3, (0.025,) * 3, wl_tf, (0, 1), 0.1)
    pts = rng.uniform(-0.2, 1.2, size=(20000, 3)).astype(np.float32)  # incl. clamp region
    pts[:8] = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1, 0, 0], [0, 1, 0], [0, 0, 1],
               [-5, 7, 0.3], [0.999, 0.001, 0.5]]
    g = vrt.sample_field(gpu_ctx, dv, pts)
    o = ov.sample(pts)
    assert np.max(np.abs(g - o)) <= 2e-6


@pytest.mark.parametrize("n,dims_grid", [(48, (3, 3, 3)), (50, (4, 3, 2))])
def test_majorants_bit_exact(pkg, oracle, gpu_ctx, n, dims_grid):
    vrt = pkg.vrt
    vox = vrt.gen_fbm((n, n - 5, n - 11), seed=3)
    tf = pkg.workloads.albedo_tf(0.8, 64)
    dv = vrt.DeviceVolume(gpu_ctx, vox, (-0.5,) * 3, (0.02,) * 3, tf, (0, 1), 0.1,
                          grid_dims=dims_grid)
    ov = oracle.OracleVolume(vox, (-0.5,) * 3, (0.02,) * 3, tf, (0, 1), 0.1)
    np.testing.assert_array_equal(dv.majorant_array(), ov.build_majorants(dims_grid))


def _raycast_pair(pkg, oracle, ctx, wl):
    gs = pkg.workloads.GpuScene(ctx, wl)
    gs.render()
    ctx.sync()
    g = gs.frame.color_array()
    o = oracle.OracleWorkload(wl).raycast()
    return g, o


def test_raycast_small(pkg, oracle, gpu_ctx):
    wl = pkg.workloads.blobs_raycast(n=64, width=128, height=96)
    g, o = _raycast_pair(pkg, oracle, gpu_ctx, wl)
    assert np.max(np.abs(g - o)) <= 1e-3


def test_raycast_blobs_128_512(pkg, oracle, gpu_ctx):
    """The raycast benchmark configuration itself: 128^3 blobs, 512x512."""
    wl = pkg.workloads.blobs_raycast()
    g, o = _raycast_pair(pkg, oracle, gpu_ctx, wl)
    err = np.abs(g - o)
    print("raycast 512^2 max-abs", err.max(), "mean", err.mean())
    assert err.max() <= 1e-3


def test_raycast_uint8_srgb(pkg, oracle, gpu_ctx):
    vrt = pkg.vrt
    wl = pkg.workloads.blobs_raycast(n=32, width=64, height=64)
    ref = oracle.OracleWorkload(wl).raycast()
    for ct, fn in ((vrt.TYPE_UFIXED8_VEC4, oracle.lib().orc_cvt_uint32),
                   (vrt.TYPE_UFIXED8_RGBA_SRGB, oracle.lib().orc_cvt_uint32_srgb)):
        gs = pkg.workloads.GpuScene(gpu_ctx, wl, color_type=ct)
        gs.render()
        gpu_ctx.sync()
        g = gs.frame.color_array()
        exp = np.array([fn(np.ascontiguousarray(p).ctypes.data) for p in ref.reshape(-1, 4)],
                       np.uint32).reshape(g.shape)
        gb = g.view(np.uint8).reshape(*g.shape, 4).astype(int)
        eb = exp.view(np.uint8).reshape(*g.shape, 4).astype(int)
        # 8-bit truncation of values that agree within 1e-3 differs by <= 1 LSB
        assert np.max(np.abs(gb - eb)) <= 1


def _pt_workload(pkg, n=32, w=64, h=48):
    return pkg.workloads.fbm_pathtrace(n=n, width=w, height=h, frames=1, brick=8)


def test_delta_track_stream_aligned(pkg, oracle, gpu_ctx):
    vrt = pkg.vrt
    wl = _pt_workload(pkg)
    dv = _gpu_volume(pkg, gpu_ctx, wl)
    scene = vrt.make_scene([dv])
    ow = oracle.OracleWorkload(wl)
    rng = np.random.default_rng(11)
    n = 4096
    ori = rng.uniform(-1.5, 1.5, (n, 3))
    ori /= np.maximum(1.0, np.abs(ori).max(axis=1, keepdims=True) / 1.5)
    target = rng.uniform(-0.4, 0.4, (n, 3))
    d = target - ori
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([ori, d, np.zeros((n, 1)), np.full((n, 1), 1e30)], 1).astype(np.float32)
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    tg, sg = vrt.delta_track(gpu_ctx, scene, rays, seeds)
    import ctypes as C
    to = np.empty(n, np.float32)
    so = seeds.copy()
    for i in range(n):
        st = C.c_uint32(int(seeds[i]))
        t = C.c_float()
        hit = oracle.lib().orc_delta_track(C.byref(ow.vol.s), rays[i, :3].ctypes.data,
                                           rays[i, 3:6].ctypes.data, float(rays[i, 6]),
                                           float(rays[i, 7]), C.byref(st), C.byref(t), 0)
        to[i] = t.value if hit else np.inf
        so[i] = st.value
    same = (np.isinf(tg) & np.isinf(to)) | (np.abs(tg - to) <= 1e-4 * np.maximum(1, np.abs(to)))
    same &= sg == so
    print("delta-track stream agreement", same.mean())
    assert same.mean() >= 0.99


def test_delta_track_homogeneous_closed_form(pkg, gpu_ctx):
    """Constant density: P(no collision over length L) = exp(-sigma L)."""
    vrt = pkg.vrt
    vox = np.full((16, 16, 16), 0.5, np.float32)
    tf = pkg.workloads.albedo_tf(0.8, 256)  # alpha = value -> 0.5
    ud = 0.5
    dv = vrt.DeviceVolume(gpu_ctx, vox, (-0.5,) * 3, (1 / 15,) * 3, tf, (0, 1), ud,
                          grid_dims=(2, 2, 2))
    scene = vrt.make_scene([dv])
    n = 1 << 18
    rng = np.random.default_rng(5)
    xy = rng.uniform(-0.4, 0.4, (n, 2))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:2] = xy
    rays[:, 2] = -2.0
    rays[:, 5] = 1.0
    rays[:, 7] = 1e30
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    t, _ = vrt.delta_track(gpu_ctx, scene, rays, seeds)
    alpha = tf[127, 3] + (tf[128, 3] - tf[127, 3]) * (0.5 * 256 - 0.5 - 127)
    sigma = alpha / ud
    p_exp = math.exp(-sigma * 1.0)
    p_hat = np.isinf(t).mean()
    tol = 5 * math.sqrt(p_exp * (1 - p_exp) / n)
    assert abs(p_hat - p_exp) <= tol, (p_hat, p_exp)
    # conditional mean depth of collisions (truncated exponential), from z = -0.5
    d = t[np.isfinite(t)] - 1.5
    m_exp = 1 / sigma - math.exp(-sigma) / (1 - math.exp(-sigma))
    assert abs(d.mean() - m_exp) <= 5 * d.std() / math.sqrt(d.size)


def test_pathtrace_stream_aligned(pkg, oracle, gpu_ctx):
    wl = _pt_workload(pkg)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    ow = oracle.OracleWorkload(wl)
    agree = []
    for accum_id in range(4):
        gs.frame.reset()
        gs.render(accum_id)
        gpu_ctx.sync()
        g = gs.frame.color_array()
        o, _ = ow.pathtrace_frame(accum_id)
        # after a reset the running mean holds alpha * sample, alpha = 1/(accum_id+1)
        o = (np.float32(1) / np.float32(accum_id + 1)) * o
        agree.append(oracle.stream_agreement(g, o))
    print("per-pixel stream agreement", agree)
    assert min(agree) >= oracle.STREAM_AGREEMENT_MIN


def test_pathtrace_mean_parity(pkg, oracle, gpu_ctx):
    """64 accumulated frames: GPU running mean vs the oracle's FP32 running mean
    (same RNG streams) and vs its FP64 restatement."""
    wl = _pt_workload(pkg)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, accum64=True)
    frames = 64
    gs.frame.reset()
    for a in range(frames):
        gs.render(a)
    gpu_ctx.sync()
    g = gs.frame.accum_array()
    g64 = gs.frame.accum64_array() / frames
    ow = oracle.OracleWorkload(wl)
    o32 = ow.pathtrace_accumulate(frames)
    o64 = ow.pathtrace_accumulate(frames, fp64=True)
    e32, e64 = relmse(g, o32), relmse(g, o64)
    print("relMSE vs oracle fp32", e32, "vs fp64", e64)
    assert e32 <= 1e-3 and e64 <= 1e-3
    # the float running mean equals the double sum / n to float rounding
    assert np.max(np.abs(g[..., :3] - g64[..., :3]) / np.maximum(1e-3, np.abs(g64[..., :3]))) < 1e-4


def test_pathtrace_anisotropic_phase(pkg, oracle, gpu_ctx):
    """Henyey-Greenstein g = 0.3 (the extension the fBm workload uses):
    stream-aligned with the oracle and within relMSE 1e-3 of its FP32 and FP64
    means; the image differs from the isotropic one."""
    import dataclasses
    wl = _pt_workload(pkg)
    wl = dataclasses.replace(wl, params={**wl.params, "phase_g": 0.3})
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    ow = oracle.OracleWorkload(wl)
    gs.frame.reset()
    gs.render(0)
    gpu_ctx.sync()
    agree = oracle.stream_agreement(gs.frame.color_array(), ow.pathtrace_frame(0)[0])
    print("HG stream agreement", agree)
    assert agree >= oracle.STREAM_AGREEMENT_MIN, agree
    frames = 64
    gs.frame.reset()
    for a in range(frames):
        gs.render(a)
    gpu_ctx.sync()
    g = gs.frame.accum_array()
    e32 = relmse(g, ow.pathtrace_accumulate(frames))
    e64 = relmse(g, ow.pathtrace_accumulate(frames, fp64=True))
    print("HG relMSE vs oracle fp32", e32, "fp64", e64)
    assert e32 <= 1e-3 and e64 <= 1e-3
    iso = oracle.OracleWorkload(_pt_workload(pkg)).pathtrace_accumulate(frames)
    assert relmse(g, iso) > 10 * max(e32, 1e-6)


def test_pathtrace_counts_lookups(pkg, oracle, gpu_ctx):
    wl = _pt_workload(pkg)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    gs.frame.reset()
    gs.render(0)
    gpu_ctx.sync()
    lookups = int(gs.frame.counter_values()[0])
    _, lo = oracle.OracleWorkload(wl).pathtrace_frame(0)
    assert lookups > 0
    assert abs(lookups - lo) <= 0.02 * lo


def test_clip_planes_parity(pkg, oracle, gpu_ctx):
    """Two oblique clip planes cutting the volume: raycast within 1e-3 of the
    oracle, path tracer stream-aligned with it (planes clip every pass)."""
    import dataclasses
    planes = [(0.3, 0.2, 1.0, 0.05), (-1.0, 0.0, 0.25, 0.2)]
    wl = pkg.workloads.blobs_raycast(n=64, width=128, height=96)
    wl = dataclasses.replace(wl, params={**wl.params, "clip_planes": planes})
    g, o = _raycast_pair(pkg, oracle, gpu_ctx, wl)
    assert np.max(np.abs(g - o)) <= 1e-3
    full, _ = _raycast_pair(pkg, oracle, gpu_ctx, pkg.workloads.blobs_raycast(n=64, width=128,
                                                                              height=96))
    assert np.max(np.abs(full - g)) > 0.05  # the planes remove part of the volume

    wl = _pt_workload(pkg)
    wl = dataclasses.replace(wl, params={**wl.params, "clip_planes": planes})
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    ow = oracle.OracleWorkload(wl)
    for accum_id in range(2):
        gs.frame.reset()
        gs.render(accum_id)
        gpu_ctx.sync()
        g = gs.frame.color_array()
        o = (np.float32(1) / np.float32(accum_id + 1)) * ow.pathtrace_frame(accum_id)[0]
        agree = oracle.stream_agreement(g, o)
        print("clip planes stream agreement", accum_id, agree)
        assert agree >= oracle.STREAM_AGREEMENT_MIN, agree


def test_thin_lens_parity(pkg, oracle, gpu_ctx):
    """A thin-lens camera (apertureRadius / focusDistance): raycast within
    1e-3 of the oracle and unlike the pinhole image, path tracer
    stream-aligned with the oracle (the lens sample drawn before the shutter
    draw)."""
    import dataclasses
    wl = pkg.workloads.blobs_raycast(n=64, width=128, height=96)
    lens = {"aperture_radius": 0.08, "focus_distance": 1.6}
    wl_lens = dataclasses.replace(wl, camera={**wl.camera, **lens})
    g, o = _raycast_pair(pkg, oracle, gpu_ctx, wl_lens)
    assert np.max(np.abs(g - o)) <= 1e-3
    pin, _ = _raycast_pair(pkg, oracle, gpu_ctx, wl)
    assert np.max(np.abs(pin - g)) > 0.05  # the lens blurs the volume out of focus

    wl = _pt_workload(pkg)
    wl = dataclasses.replace(wl, camera={**wl.camera, **lens})
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    ow = oracle.OracleWorkload(wl)
    for accum_id in range(2):
        gs.frame.reset()
        gs.render(accum_id)
        gpu_ctx.sync()
        g = gs.frame.color_array()
        o = (np.float32(1) / np.float32(accum_id + 1)) * ow.pathtrace_frame(accum_id)[0]
        agree = oracle.stream_agreement(g, o)
        print("thin lens stream agreement", accum_id, agree)
        assert agree >= oracle.STREAM_AGREEMENT_MIN, agree


def test_camera_lens_refused(pkg, gpu_ctx):
    """Invalid lens parameters, and a lens with the radiance cache (whose
    kernels compile the lens out), are errors."""
    import dataclasses
    wl = _pt_workload(pkg, n=16, w=16, h=16)
    for bad in ({"aperture_radius": -0.1}, {"aperture_radius": 0.1, "focus_distance": 0.0}):
        gs = pkg.workloads.GpuScene(gpu_ctx, dataclasses.replace(wl, camera={**wl.camera, **bad}))
        with pytest.raises(pkg.vrt.VrtError, match="aperture_radius|focus_distance"):
            gs.render(0)
    gs = pkg.workloads.GpuScene(gpu_ctx, dataclasses.replace(
        wl, camera={**wl.camera, "aperture_radius": 0.05, "focus_distance": 2.0}))
    nrc = pkg.vrt.Nrc(gpu_ctx, max_batch=16384)
    try:
        with pytest.raises(pkg.vrt.VrtError, match="thin-lens"):
            gs.render_nrc(nrc, gs.nrc_options(train=False), 0)
    finally:
        nrc.close()


@pytest.mark.parametrize("field", ["fbm", "tets"])
def test_pathtrace_ratio_tracking_parity(pkg, oracle, gpu_ctx, field):
    """Ratio-tracked shadow and AO rays (the ratio_tracking extension): the GPU
    frames stream-aligned with the oracle's, and the 32-frame mean within
    relMSE 1e-3 of the oracle's FP32 and FP64 means."""
    import dataclasses
    if field == "tets":
        wl = pkg.workloads.tets_pathtrace(n=8, width=64, height=48, frames=1, grid=16)
    else:
        wl = _pt_workload(pkg)
    wl = dataclasses.replace(wl, params=dict(wl.params, ratio_tracking=True))
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    ow = oracle.OracleWorkload(wl)
    agree = []
    for accum_id in range(3):
        gs.frame.reset()
        gs.render(accum_id)
        gpu_ctx.sync()
        o, _ = ow.pathtrace_frame(accum_id)
        o = (np.float32(1) / np.float32(accum_id + 1)) * o
        agree.append(oracle.stream_agreement(gs.frame.color_array(), o))
    frames = 32
    gs.frame.reset()
    for a in range(frames):
        gs.render(a)
    gpu_ctx.sync()
    g = gs.frame.accum_array()
    e32 = relmse(g, ow.pathtrace_accumulate(frames))
    e64 = relmse(g, ow.pathtrace_accumulate(frames, fp64=True))
    print(field, "ratio tracking: stream agreement", agree, "relMSE fp32", e32, "fp64", e64)
    assert min(agree) >= oracle.STREAM_AGREEMENT_MIN
    assert e32 <= 1e-3 and e64 <= 1e-3
