"""Unstructured (tetrahedral) field in the oracle (C4): intersectTet point
location and interpolation, cell handling, the running cell size and the
majorant grid, each against closed forms, brute force or an independent
float64 computation (UnstructuredField.cpp:29-198, 376-399,
DeviceCopyableObjects.h:128-164, 482-527, Plane.h:56-96)."""
import ctypes as C
import math

import numpy as np
import pytest


def _lattice(pkg, n, jitter=0.15, seed=5):
    return pkg.vrt.gen_tet_lattice(n, jitter, 32, seed)


def _field(oracle, pos, idx, vals=None, tf=None, **kw):
    if tf is None:
        tf = np.stack([np.linspace(0, 1, 16)] * 4, 1).astype(np.float32)
    return oracle.OracleVolume.from_unstructured(pos, idx, tf, (0.0, 1.0), 0.1, vals, **kw)


def _barycentric_f64(pos, idx, p):
    """Containing tetrahedra of points p by float64 barycentric coordinates:
    (b (P, T, 4), inside (P, T) with every coordinate >= -tol)."""
    v = pos[idx].astype(np.float64)  # (T, 4, 3)
    M = np.transpose(v[:, 1:] - v[:, :1], (0, 2, 1))  # (T, 3, 3) columns e1 e2 e3
    Minv = np.linalg.inv(M)
    r = p[:, None, :].astype(np.float64) - v[None, :, 0, :]  # (P, T, 3)
    b123 = np.einsum("tij,ptj->pti", Minv, r)
    b0 = 1.0 - b123.sum(-1)
    return np.concatenate([b0[..., None], b123], -1)


def test_generator_lattice(pkg):
    """Kuhn split: 6 tetrahedra per cube, all of the undistorted volume
    h^3/6 at zero jitter, filling the cube; boundary vertices stay on it."""
    n = 5
    pos, val, idx = _lattice(pkg, n, jitter=0.0)
    assert pos.shape == ((n + 1) ** 3, 3) and idx.shape == (6 * n ** 3, 4)
    v = pos[idx].astype(np.float64)
    vol = np.abs(np.linalg.det(v[:, 1:] - v[:, :1])) / 6
    np.testing.assert_allclose(vol, 1 / (6 * n ** 3), rtol=1e-5)
    assert abs(vol.sum() - 1.0) < 1e-5
    pos, _, idx = _lattice(pkg, n, jitter=0.15)
    on_face = np.isclose(np.abs(pos), 0.5)
    assert np.all(np.abs(pos) <= 0.5 + 1e-7) and on_face.any()
    v = pos[idx].astype(np.float64)
    assert abs(np.abs(np.linalg.det(v[:, 1:] - v[:, :1])).sum() / 6 - 1.0) < 1e-5
    assert 0.0 <= val.min() and val.max() <= 1.0 and val.std() > 0.01


def test_affine_data_reproduced(pkg, oracle):
    """Barycentric interpolation reproduces an affine vertex field at every
    point inside the mesh (the mesh fills [-0.5, 0.5]^3)."""
    pos, _, idx = _lattice(pkg, 6)
    aff = lambda p: 0.3 + 0.2 * p[..., 0] - 0.1 * p[..., 1] + 0.25 * p[..., 2]
    vals = aff(pos.astype(np.float64)).astype(np.float32)
    f = _field(oracle, pos, idx, vals)
    p = np.random.default_rng(1).uniform(-0.499, 0.499, (20000, 3)).astype(np.float32)
    got = f.sample(p)
    assert not np.isnan(got).any()
    np.testing.assert_allclose(got, aff(p.astype(np.float64)), atol=2e-6)


def test_point_location_matches_brute_force_and_f64(pkg, oracle):
    """The oracle's grid finds the same cell as testing every cell in id
    order; membership and value agree with float64 barycentric coordinates
    away from faces (|b| > 1e-5); points outside the mesh give NaN."""
    pos, val, idx = _lattice(pkg, 5)
    f = _field(oracle, pos, idx, val)
    p = np.random.default_rng(2).uniform(-0.6, 0.6, (3000, 3)).astype(np.float32)
    got, brute = f.sample(p), f.sample_brute(p)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(brute))
    ok = ~np.isnan(got)
    np.testing.assert_array_equal(got[ok], brute[ok])
    b = _barycentric_f64(pos, idx, p)
    inside = (b >= 0).all(-1)
    clear = (np.abs(b) > 1e-5).all(-1)  # not within 1e-5 of a face plane
    in_any = inside.any(-1)
    sure = clear.all(-1) | ~in_any
    np.testing.assert_array_equal(ok[sure], in_any[sure])
    t = np.argmax(inside, -1)
    ref = (b[np.arange(len(p)), t] * val[idx[t]].astype(np.float64)).sum(-1)
    m = ok & in_any & sure
    np.testing.assert_allclose(got[m], ref[m], atol=1e-5)
    assert 0.3 < ok.mean() < 0.9  # the box [-0.6, 0.6]^3 is half inside


def test_cell_data_without_vertex_data(pkg, oracle):
    """No vertex data: the value inside a cell is its cell.data value."""
    pos, _, idx = _lattice(pkg, 4)
    cd = (np.arange(len(idx)) * 1e-3).astype(np.float32)
    f = _field(oracle, pos, idx, None, cell_data=cd)
    cent = pos[idx].astype(np.float64).mean(1).astype(np.float32)
    np.testing.assert_array_equal(f.sample(cent), cd)


def test_cell_size_running_average(pkg, oracle):
    """Every cell of the unjittered lattice has the cube as its box, so the
    reference's running sum of diagonal / (cells so far) is h sqrt(3) H_T
    (H_T the T-th harmonic number), not the mean diagonal h sqrt(3)."""
    n = 4
    pos, val, idx = _lattice(pkg, n, jitter=0.0)
    f = _field(oracle, pos, idx, val)
    bounds, cs, kept, skipped = f.unstructured_info()
    T = len(idx)
    assert kept == T and skipped == 0
    np.testing.assert_allclose(bounds, [-0.5] * 3 + [0.5] * 3, atol=1e-7)
    H = sum(1.0 / k for k in range(1, T + 1))
    np.testing.assert_allclose(cs, math.sqrt(3) / n * H, rtol=1e-4)


def test_cells_skipped_or_rejected(pkg, oracle):
    """Cells with fewer than 4 indices or a type/count mismatch are skipped
    (UnstructuredField.cpp:135-147); a wedge is not restated."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [1, 1, 0]], np.float32)
    idx = np.array([0, 1, 2, 3, 0, 1, 2, 1, 2, 3, 4, 0, 1, 2, 4], np.uint32)
    ci = np.array([0, 4, 7, 11], np.uint32)
    ct = np.array([10, 10, 12, 0], np.uint8)  # tet, 3 indices, "hex" with 4, tet
    f = _field(oracle, pos, idx, np.linspace(0, 1, 6).astype(np.float32), cell_index=ci,
               cell_type=ct)
    bounds, cs, kept, skipped = f.unstructured_info()
    assert (kept, skipped) == (2, 2)
    w = _field(oracle, pos, np.arange(6, dtype=np.uint32), np.zeros(6, np.float32),
               cell_index=np.array([0], np.uint32), cell_type=np.array([13], np.uint8))
    assert w.unstructured_info() is None


def test_majorants_bound_every_sample(pkg, oracle):
    pos, val, idx = _lattice(pkg, 6)
    tf = pkg.workloads.albedo_tf(0.8, 32)
    tf[:, 3] = np.sin(np.linspace(0, 7, 32)) ** 2  # non-monotone opacity
    f = _field(oracle, pos, idx, val, tf=tf)
    G = (5, 4, 3)
    maj = f.build_majorants(G)
    bounds = f.unstructured_info()[0]
    p = np.random.default_rng(3).uniform(bounds[:3], bounds[3:], (100000, 3)).astype(np.float32)
    s = f.sample(p)
    ok = ~np.isnan(s)
    a = f.classify(s[ok])[:, 3]
    q = (p[ok] - bounds[:3]) / (bounds[3:] - bounds[:3]) * np.array(G, np.float32)
    cell = np.clip(q.astype(int), 0, np.array(G) - 1)
    m = maj[cell[:, 2], cell[:, 1], cell[:, 0]]
    assert np.all(a <= m + 1e-6)
    assert np.mean(a >= 0.9 * m) > 0.001


@pytest.mark.parametrize("fp64", [0, 1])
def test_delta_tracking_through_gap(pkg, oracle, fp64):
    """Constant density in two slabs of cells with an empty gap between
    them (x in [-0.25, 0.25]): a ray along x escapes with probability
    exp(-sigma * 0.5); points in the gap have no value and never collide."""
    n = 8
    pos, _, idx = _lattice(pkg, n, jitter=0.0)
    cube_x = (np.arange(len(idx)) // 6) % n
    keep = (cube_x < 2) | (cube_x >= 6)
    tf = np.stack([np.full(4, 0.8)] * 3 + [np.full(4, 0.3)], 1).astype(np.float32)
    ud = 0.1
    f = oracle.OracleVolume.from_unstructured(pos, idx[keep], tf, (0, 1), ud,
                                             np.full(len(pos), 0.5, np.float32))
    f.attach_majorants((4, 2, 2))
    sigma = 0.3 / ud
    N = 30000
    rng = np.random.default_rng(5)
    st, t = C.c_uint32(0), C.c_float()
    d = np.array([1, 0, 0], np.float32)
    esc, hits = 0, []
    for i in range(N):
        o = np.array([-2.0, rng.uniform(-0.45, 0.45), rng.uniform(-0.45, 0.45)], np.float32)
        st.value = int(oracle.lib().orc_tea_seed(i, 23))
        if oracle.lib().orc_delta_track(C.byref(f.s), o.ctypes.data, d.ctypes.data, 0.0, 1e30,
                                        C.byref(st), C.byref(t), fp64):
            hits.append(t.value - 2.0)  # x of the collision
        else:
            esc += 1
    p = math.exp(-sigma * 0.5)
    assert abs(esc / N - p) <= 5 * math.sqrt(p * (1 - p) / N)
    x = np.array(hits)
    assert not np.any((x > -0.25 + 1e-4) & (x < 0.25 - 1e-4))
    assert np.all(np.abs(x) <= 0.5 + 1e-4)
