"""GPU parity of the 'nanovdb' field (C3, DeviceCopyableObjects.h:554-569)
against the oracle's NanoVDB tree walk: fetches bit-exact (same float
operations, explicit fmaf map, no contraction), majorants bit-exact, the ray
caster within 1e-3, the path tracer stream-aligned and within relMSE 1e-3 of
the oracle's mean."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tiled_grid(oracle):
    ref = oracle.ref_nvdb()
    if ref is None:
        return None
    ptr, n = C.c_void_p(), C.c_uint64()
    assert ref.ref_nvdb_make_tiled(5, C.byref(ptr), C.byref(n)) == 0
    raw = np.empty(n.value + 64, np.uint8)
    off = (-raw.ctypes.data) % 64
    buf = raw[off:off + n.value]
    C.memmove(buf.ctypes.data, ptr.value, n.value)
    ref.ref_nvdb_free(ptr)
    return buf


def _volume(pkg, ctx, buf, filter="linear", grid_dims=None):
    tf = pkg.workloads.albedo_tf(0.8, 64)
    return pkg.vrt.DeviceVolume.from_nanovdb(ctx, buf, tf, (0, 1), 1.0, grid_dims, filter=filter)


def test_nanovdb_fetch_bit_exact(pkg, oracle, gpu_ctx):
    vrt = pkg.vrt
    rng = np.random.default_rng(2)
    dense = np.zeros((30, 21, 26), np.float32)
    m = rng.random(dense.shape) < 0.1
    dense[m] = rng.uniform(0.05, 1, m.sum())
    grids = [(vrt.gen_nanovdb_from_dense(dense, 0.0, 0.25, (-2.0, 1.0, 0.5)), 0.25,
              np.array([-2.0, 1.0, 0.5]), (-3, -3, -3), (29, 24, 33))]
    tiled = _tiled_grid(oracle)
    if tiled is not None:
        grids.append((tiled, 0.5, np.array([-3.0, 1.0, 2.0]), (-14, -14, -14), (22, 22, 22)))
        grids.append((tiled, 0.5, np.array([-3.0, 1.0, 2.0]), (100, -4, -4), (140, 10, 10)))
    for buf, voxel, trans, lo, hi in grids:
        idx = rng.uniform(lo, hi, (20000, 3))
        pts = (voxel * idx + trans).astype(np.float32)
        for filt in ("linear", "nearest"):
            dv = _volume(pkg, gpu_ctx, buf, filt)
            g = vrt.sample_field(gpu_ctx, dv, pts)
            o = oracle.nvdb_sample(buf, pts, 1 if filt == "linear" else 0)
            np.testing.assert_array_equal(g, o)


def test_nanovdb_majorants_bit_exact(pkg, oracle, gpu_ctx):
    buf, _ = pkg.vrt.gen_nanovdb_cloud(n=64, seed=4)
    tf = pkg.workloads.albedo_tf(0.8, 64)
    for grid in ((5, 5, 5), (7, 4, 3)):
        dv = pkg.vrt.DeviceVolume.from_nanovdb(gpu_ctx, buf, tf, (0, 1), 1.0, grid)
        ov = oracle.OracleVolume.from_nanovdb(buf, tf, (0, 1), 1.0)
        m = ov.build_majorants(grid)
        np.testing.assert_array_equal(dv.majorant_array(), m)
        assert m.max() > 0


def _cloud(pkg, **kw):
    return pkg.workloads.cloud_nanovdb_pathtrace(n=64, width=64, height=36, frames=1, brick=8,
                                                 **kw)


def test_nanovdb_raycast(pkg, oracle, gpu_ctx):
    wl = _cloud(pkg)
    wl = dataclasses.replace(wl, renderer="raycast",
                             params=dict(background=(0, 0, 0, 1), volume_sampling_rate=2.0),
                             tf=pkg.workloads.ramp_tf(256))
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    gs.render(0)
    gpu_ctx.sync()
    g = gs.frame.color_array()
    o = oracle.OracleWorkload(wl).raycast()
    assert np.max(np.abs(g - o)) <= 1e-3
    assert (o[..., 3] > 0.05).mean() > 0.05  # the cloud covers part of the image


def test_nanovdb_pathtrace_parity(pkg, oracle, gpu_ctx):
    wl = _cloud(pkg)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, accum64=True)
    ow = oracle.OracleWorkload(wl)
    for accum_id in range(2):
        gs.frame.reset()
        gs.render(accum_id)
        gpu_ctx.sync()
        g = gs.frame.color_array()
        o = (np.float32(1) / np.float32(accum_id + 1)) * ow.pathtrace_frame(accum_id)[0]
        agree = oracle.stream_agreement(g, o)
        print("nanovdb stream agreement", accum_id, agree)
        assert agree >= oracle.STREAM_AGREEMENT_MIN, agree
    frames = 32
    gs.frame.reset()
    for a in range(frames):
        gs.render(a)
    gpu_ctx.sync()
    g = gs.frame.accum_array()
    o32 = ow.pathtrace_accumulate(frames)
    o64 = ow.pathtrace_accumulate(frames, fp64=True)
    x, eps = np.asarray(g, np.float64)[..., :3], 1e-2
    for r in (o32, o64):
        r = np.asarray(r, np.float64)[..., :3]
        e = float(np.mean((x - r) ** 2 / (r * r + eps)))
        assert e <= 1e-3, e
