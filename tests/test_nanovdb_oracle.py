"""NanoVDB grids (C3 field, DeviceCopyableObjects.h:554-569): our writer's
buffers, the oracle's tree walk and samplers, each pinned against the
reference's vendored NanoVDB 32.7 compiled into oracle/_ref (the golden
source the reference itself links), plus the trilinear closed form."""
import ctypes as C

import numpy as np
import pytest


@pytest.fixture(scope="module")
def ref(oracle):
    r = oracle.ref_nvdb()
    if r is None:
        pytest.skip("oracle/_ref/libnvdb_ref.so not built (reference tree absent)")
    return r


def _check(ref, buf):
    err = C.create_string_buffer(512)
    rc = ref.ref_nvdb_check(buf.ctypes.data, err, 512)
    assert rc == 0, err.value.decode()


def _ref_get(ref, buf, ijk):
    ijk = np.ascontiguousarray(ijk, np.int32)
    out = np.empty(len(ijk), np.float32)
    ref.ref_nvdb_get(buf.ctypes.data, len(ijk), ijk.ctypes.data, out.ctypes.data)
    return out


def _ref_sample(ref, buf, xyz, order):
    xyz = np.ascontiguousarray(xyz, np.float32)
    out = np.empty(len(xyz), np.float32)
    ref.ref_nvdb_sample(buf.ctypes.data, len(xyz), xyz.ctypes.data, order, out.ctypes.data)
    return out


def _lattice(lo, hi):
    g = np.mgrid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]].reshape(3, -1).T
    return np.ascontiguousarray(g, np.int32)


def test_writer_dense_roundtrip(pkg, oracle, ref):
    """A sparse random array written by our NanoVDB writer passes NanoVDB's
    full structural check, and NanoVDB reads back every voxel (and the
    background around it) exactly; the oracle's walk agrees."""
    rng = np.random.default_rng(7)
    dense = np.zeros((20, 27, 35), np.float32)
    mask = rng.random(dense.shape) < 0.07
    dense[mask] = rng.uniform(0.1, 1.0, mask.sum()).astype(np.float32)
    dense[3:9, 4:12, 18:26] = 0.5  # a filled block across leaf borders
    for voxel, trans in ((1.0, (0, 0, 0)), (0.25, (-2.0, 1.5, 3.0))):
        buf = pkg.vrt.gen_nanovdb_from_dense(dense, 0.0, voxel, trans)
        _check(ref, buf)
        ijk = _lattice((-3, -3, -3), (dense.shape[2] + 3, dense.shape[1] + 3, dense.shape[0] + 3))
        exp = np.zeros(len(ijk), np.float32)
        inside = np.all((ijk >= 0) & (ijk < np.array(dense.shape[::-1])), axis=1)
        exp[inside] = dense[ijk[inside, 2], ijk[inside, 1], ijk[inside, 0]]
        np.testing.assert_array_equal(_ref_get(ref, buf, ijk), exp)
        np.testing.assert_array_equal(oracle.nvdb_get(buf, ijk), exp)
        # world bbox = map([min, max + 1]) of the active voxels
        act = np.argwhere(dense != 0)[:, ::-1]
        bb = np.concatenate([act.min(0) * voxel + trans, (act.max(0) + 1) * voxel + trans])
        np.testing.assert_allclose(oracle.nvdb_world_bbox(buf), bb, rtol=0, atol=1e-12)
        wb = np.empty(6)
        ref.ref_nvdb_world_bbox(buf.ctypes.data, wb.ctypes.data)
        np.testing.assert_array_equal(wb, oracle.nvdb_world_bbox(buf))


def test_tiled_grid_walk_matches_nanovdb(oracle, ref):
    """A grid built by NanoVDB's own builder with root, upper and lower tiles
    (active and inactive), negative coordinates and a scaled, translated map:
    the oracle's walk equals NanoVDB's getValue on a lattice covering every
    kind of node, and both samplers agree bit for bit at random points."""
    ptr, n = C.c_void_p(), C.c_uint64()
    assert ref.ref_nvdb_make_tiled(5, C.byref(ptr), C.byref(n)) == 0
    raw = np.empty(n.value + 64, np.uint8)
    off = (-raw.ctypes.data) % 64
    buf = raw[off:off + n.value]
    C.memmove(buf.ctypes.data, ptr.value, n.value)
    ref.ref_nvdb_free(ptr)
    _check(ref, buf)
    for lo, hi in (((-16, -16, -16), (24, 24, 24)), ((100, -4, -4), (160, 10, 10)),
                   ((4080, 0, -2), (4110, 6, 6))):
        ijk = _lattice(lo, hi)
        got, exp = oracle.nvdb_get(buf, ijk), _ref_get(ref, buf, ijk)
        np.testing.assert_array_equal(got, exp)
    vals = set(np.unique(_ref_get(ref, buf, _lattice((-16, -16, -16), (24, 24, 24)))))
    assert {np.float32(0.25), np.float32(0.9)} <= vals  # background and the lower tile
    rng = np.random.default_rng(3)
    # world = 0.5 * index + (-3, 1, 2)
    for lo, hi in (((-10, -10, -10), (22, 22, 22)), ((118, -2, -2), (140, 8, 8)),
                   ((4086, 0, 0), (4102, 6, 6))):
        idx = rng.uniform(lo, hi, (4000, 3))
        xyz = (0.5 * idx + np.array([-3.0, 1.0, 2.0])).astype(np.float32)
        for order in (0, 1):
            np.testing.assert_array_equal(oracle.nvdb_sample(buf, xyz, order),
                                          _ref_sample(ref, buf, xyz, order))


def test_trilinear_closed_form(pkg, oracle):
    """NanoVDB values sit at integer index positions (no half-voxel shift):
    an affine field is reproduced exactly by the linear sampler inside the
    grid, and the nearest sampler returns the voxel of the rounded index."""
    g = np.mgrid[0:12, 0:10, 0:9].astype(np.float64)  # z, y, x
    dense = (0.1 + 0.02 * g[2] + 0.03 * g[1] + 0.05 * g[0]).astype(np.float32)
    buf = pkg.vrt.gen_nanovdb_from_dense(dense, 0.0, 1.0, (0, 0, 0))
    rng = np.random.default_rng(1)
    xyz = rng.uniform([0, 0, 0], [8, 9, 11], (3000, 3)).astype(np.float32)
    lin = oracle.nvdb_sample(buf, xyz, 1)
    exp = 0.1 + 0.02 * xyz[:, 0] + 0.03 * xyz[:, 1] + 0.05 * xyz[:, 2]
    np.testing.assert_allclose(lin, exp, atol=2e-6)
    near = oracle.nvdb_sample(buf, xyz, 0)
    r = np.rint(xyz).astype(int)
    np.testing.assert_array_equal(near, dense[r[:, 2], r[:, 1], r[:, 0]])


def test_cloud_generator(pkg, oracle, ref):
    """The C3 cloud (small edge): a valid grid whose active voxels carry the
    thresholded densities in (0, 1], about the requested fraction active;
    oracle and NanoVDB samplers agree bit for bit inside its bounds."""
    buf, tau = pkg.vrt.gen_nanovdb_cloud(n=64, active_fraction=0.05, seed=3)
    _check(ref, buf)
    ijk = _lattice((0, 0, 0), (64, 64, 64))
    v = oracle.nvdb_get(buf, ijk)
    frac = float((v > 0).mean())
    assert 0.02 < frac < 0.1, frac
    assert v.max() <= 1.0 and 0 < tau < 1
    np.testing.assert_array_equal(v, _ref_get(ref, buf, ijk))
    bb = oracle.nvdb_world_bbox(buf)
    rng = np.random.default_rng(9)
    xyz = rng.uniform(bb[:3], bb[3:], (5000, 3)).astype(np.float32)
    np.testing.assert_array_equal(oracle.nvdb_sample(buf, xyz, 1), _ref_sample(ref, buf, xyz, 1))
