"""GPU parity of the unstructured (tetrahedral) field (C4,
DeviceCopyableObjects.h:482-527, Plane.h:56-96) against the oracle: field
info (bounds, the running cell size) bit-exact, point location and values
equal (same float operations; only points within rounding of a shared face
may take the neighbouring cell, whose value agrees to rounding), majorants
bit-exact, the ray caster within 1e-3, the path tracer stream-aligned and
within relMSE 1e-3 of the oracle's FP32 and FP64 means."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(pkg, ctx, n=8, jitter=0.15, keep=None):
    pos, val, idx = pkg.vrt.gen_tet_lattice(n, jitter, 32, 5)
    if keep is not None:
        idx = idx[keep(idx)]
    return pos, val, idx, pkg.vrt.DeviceUnstructured(ctx, pos, idx, vertex_data=val)


def test_info_matches_oracle(pkg, oracle, gpu_ctx):
    pos, val, idx, f = _field(pkg, gpu_ctx, n=9)
    ov = oracle.OracleVolume.from_unstructured(pos, idx, np.ones((4, 4), np.float32), (0, 1),
                                               1.0, val)
    bounds, cs, kept, skipped = ov.unstructured_info()
    info = f.info
    np.testing.assert_array_equal(np.array(list(info.bounds_lo) + list(info.bounds_hi),
                                           np.float32), bounds)
    assert np.float32(info.cell_size) == np.float32(cs)
    assert (info.num_cells, info.skipped_cells) == (kept, skipped) == (len(idx), 0)
    assert info.list_entries >= len(idx) and 1.0 <= info.expected_tests < 8.0
    print("tet grid", list(info.grid_dims), "mean list", info.mean_list, "expected tests",
          info.expected_tests, "build ms", info.build_ms)


def test_rejects_unsupported_cells(pkg, gpu_ctx):
    vrt = pkg.vrt
    pos = np.random.default_rng(0).random((8, 3)).astype(np.float32)
    with pytest.raises(vrt.VrtError, match="tetrahedral"):
        vrt.DeviceUnstructured(gpu_ctx, pos, np.arange(8, dtype=np.uint32),
                               np.array([0], np.uint32), np.array([12], np.uint8),
                               np.zeros(8, np.float32))
    with pytest.raises(vrt.VrtError, match="outside"):
        vrt.DeviceUnstructured(gpu_ctx, pos, np.array([0, 1, 2, 9], np.uint32),
                               vertex_data=np.zeros(8, np.float32))


@pytest.mark.parametrize("holes", [False, True])
def test_sample_parity(pkg, oracle, gpu_ctx, holes):
    """Values at random points inside and around the mesh (and, with holes,
    in empty cubes inside the bounds): NaN exactly where the oracle finds no
    cell, equal values elsewhere."""
    keep = (lambda idx: (np.arange(len(idx)) // 6) % 3 != 1) if holes else None
    pos, val, idx, f = _field(pkg, gpu_ctx, n=10, keep=keep)
    tf = pkg.workloads.albedo_tf(0.8, 64)
    dv = pkg.vrt.DeviceVolume.from_unstructured(gpu_ctx, f, tf, (0, 1), 1.0)
    ov = oracle.OracleVolume.from_unstructured(pos, idx, tf, (0, 1), 1.0, val)
    p = np.random.default_rng(7).uniform(-0.55, 0.55, (100000, 3)).astype(np.float32)
    g = pkg.vrt.sample_field(gpu_ctx, dv, p)
    o = ov.sample(p)
    miss_g, miss_o = np.isnan(g), np.isnan(o)
    assert (miss_g != miss_o).sum() <= 2  # only points within rounding of the mesh surface
    both = ~miss_g & ~miss_o
    assert np.mean(g[both] == o[both]) >= 0.9999
    np.testing.assert_allclose(g[both], o[both], atol=1e-6)
    assert 0.2 < both.mean() < 0.9


def test_majorants_bit_exact(pkg, oracle, gpu_ctx):
    pos, val, idx, f = _field(pkg, gpu_ctx, n=9)
    tf = pkg.workloads.albedo_tf(0.8, 32)
    tf[:, 3] = np.sin(np.linspace(0, 7, 32)) ** 2
    ov = oracle.OracleVolume.from_unstructured(pos, idx, tf, (0, 1), 1.0, val)
    for grid in ((5, 5, 5), (7, 4, 3), (16, 16, 16)):
        dv = pkg.vrt.DeviceVolume.from_unstructured(gpu_ctx, f, tf, (0, 1), 1.0, grid)
        m = ov.build_majorants(grid)
        np.testing.assert_array_equal(dv.majorant_array(), m)
        assert m.max() > 0


def test_raycast_parity(pkg, oracle, gpu_ctx):
    wl = pkg.workloads.tets_raycast(n=12, width=128, height=96)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    gs.render(0)
    gpu_ctx.sync()
    g = gs.frame.color_array()
    o = oracle.OracleWorkload(wl).raycast()
    err = np.max(np.abs(g - o))
    print("tets raycast max-abs", err)
    assert err <= 1e-3
    assert (o[..., 3] > 0.05).mean() > 0.1


def _pt(pkg, **kw):
    return pkg.workloads.tets_pathtrace(n=10, width=64, height=48, frames=1, grid=8, **kw)


def test_pathtrace_parity(pkg, oracle, gpu_ctx):
    wl = _pt(pkg)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, accum64=True)
    ow = oracle.OracleWorkload(wl)
    for accum_id in range(2):
        gs.frame.reset()
        gs.render(accum_id)
        gpu_ctx.sync()
        g = gs.frame.color_array()
        o = (np.float32(1) / np.float32(accum_id + 1)) * ow.pathtrace_frame(accum_id)[0]
        agree = oracle.stream_agreement(g, o)
        print("tets stream agreement", accum_id, agree)
        assert agree >= oracle.STREAM_AGREEMENT_MIN, agree
    frames = 32
    gs.frame.reset()
    for a in range(frames):
        gs.render(a)
    gpu_ctx.sync()
    x = np.asarray(gs.frame.accum_array(), np.float64)[..., :3]
    for fp64 in (False, True):
        r = np.asarray(ow.pathtrace_accumulate(frames, fp64=fp64), np.float64)[..., :3]
        e = float(np.mean((x - r) ** 2 / (r * r + 1e-2)))
        print("tets relMSE vs oracle", "fp64" if fp64 else "fp32", e)
        assert e <= 1e-3, (fp64, e)
    assert x.mean() > 0.01


def test_pathtrace_with_holes(pkg, oracle, gpu_ctx):
    """Cubes removed inside the bounds: delta tracking draws no acceptance
    number where no cell holds the point, on both sides."""
    wl = _pt(pkg)
    t = dict(wl.tets)
    keep = ((np.arange(len(t["index"])) // 6) // 10) % 3 != 1
    t["index"] = t["index"][keep]
    wl = dataclasses.replace(wl, tets=t)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    gs.frame.reset()
    gs.render(0)
    gpu_ctx.sync()
    agree = oracle.stream_agreement(gs.frame.color_array(),
                                    oracle.OracleWorkload(wl).pathtrace_frame(0)[0])
    print("tets with holes stream agreement", agree)
    assert agree >= oracle.STREAM_AGREEMENT_MIN, agree
