"""The oracle's instanced meshes (intersect(Ray, Instance),
DeviceCopyableObjects.h:1772-1814; Instance.cpp:88-97), pinned on the CPU:

  transforms     affine_inv is inverse(L) (numpy's inverse to float rounding),
                 trans_inv = -t, normal_xfm = transpose(affine_inv); the
                 library's vrt_instance_transforms gives the same floats;
                 singular and non-finite L are refused
  world hits     an instance is hit where its mesh, moved to world space
                 vertex by vertex, is hit: same triangle, the same world
                 distance to float rounding (object-space rays keep t)
  ties           two coincident instances: the lower index wins; a mesh and
                 a coincident instance: the mesh wins
  any hit        any-hit rays report a hit exactly where closest-hit rays do
"""
import math

import numpy as np
import pytest


def _xf(rng, scale_lo=0.3, scale_hi=1.5):
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    ang = rng.uniform(0, 2 * math.pi)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)
    L = R @ np.diag(rng.uniform(scale_lo, scale_hi, 3))
    t = rng.uniform(-1, 1, 3)
    return np.concatenate([L, t[:, None]], 1).astype(np.float32)


def _aimed(rng, n, centre, spread=0.6, dist=3.0):
    """Rays from a sphere of radius `dist` around `centre` towards points
    within `spread` of it."""
    o = rng.normal(size=(n, 3))
    o = centre + dist * o / np.linalg.norm(o, axis=1, keepdims=True)
    d = centre + rng.uniform(-spread, spread, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3:6] = d
    r[:, 7] = 1e30
    return r


def test_transforms_match_numpy_and_library(pkg, oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        xf = _xf(rng)
        a, t, n = oracle.instance_transforms(xf)
        L = xf[:, :3].astype(np.float64)
        np.testing.assert_allclose(a, np.linalg.inv(L), rtol=2e-6, atol=2e-6)
        np.testing.assert_array_equal(t, -xf[:, 3])
        np.testing.assert_array_equal(n, a.T)
        la, lt, ln = pkg.vrt.instance_transforms(xf)
        np.testing.assert_array_equal(la, a)
        np.testing.assert_array_equal(lt, t)
        np.testing.assert_array_equal(ln, n)
    sing = np.zeros((3, 4), np.float32)
    sing[0, 0] = sing[1, 1] = 1.0  # rank 2
    assert oracle.instance_transforms(sing) is None
    bad = _xf(rng)
    bad[1, 2] = np.nan
    assert oracle.instance_transforms(bad) is None
    with pytest.raises(pkg.vrt.VrtError):
        pkg.vrt.instance_transforms(sing)


def test_instance_hits_equal_world_mesh_hits(pkg, oracle):
    rng = np.random.default_rng(2)
    v, idx = pkg.vrt.gen_icosphere(level=3, radius=0.6, amp=0.1, freq=3.0, seed=5)
    for trial in range(4):
        xf = _xf(rng)
        inst = oracle.OracleInstance(oracle.OracleMesh(v, idx), xf)
        vw = (v.astype(np.float64) @ xf[:, :3].T.astype(np.float64) + xf[:, 3]).astype(np.float32)
        sc_i = oracle.OracleScene([], [], [], [inst])
        sc_w = oracle.OracleScene([], [], [oracle.OracleMesh(vw, idx)])
        r = _aimed(rng, 4000, xf[:, 3].astype(np.float64))
        ti, pi, si = oracle.scene_intersect_surfaces(sc_i, r)
        tw, pw, sw = oracle.scene_intersect_surfaces(sc_w, r)
        hit = np.isfinite(tw)
        assert hit.mean() > 0.1
        # the same rays hit, but for grazing ones the rounding of either
        # construction decides
        assert np.mean(np.isfinite(ti) == hit) > 0.995
        both = hit & np.isfinite(ti)
        assert np.mean(pi[both] == pw[both]) > 0.99
        same = both & (pi == pw)
        np.testing.assert_allclose(ti[same], tw[same], rtol=2e-4, atol=1e-6)
        assert np.all(si[np.isfinite(ti)] == 0) and np.all(si[~np.isfinite(ti)] == -1)


def test_tie_rules(pkg, oracle):
    rng = np.random.default_rng(3)
    v, idx = pkg.vrt.gen_icosphere(level=2, radius=0.5, amp=0.05, freq=2.0, seed=1)
    xf = _xf(rng, 0.8, 1.2)
    vw = (v.astype(np.float64) @ xf[:, :3].T.astype(np.float64) + xf[:, 3]).astype(np.float32)
    r = _aimed(rng, 3000, xf[:, 3].astype(np.float64))
    # two coincident instances: every hit is instance 0 (the lower index)
    sc = oracle.OracleScene([], [], [], [oracle.OracleInstance(oracle.OracleMesh(v, idx), xf),
                                        oracle.OracleInstance(oracle.OracleMesh(v, idx), xf)])
    t, p, s = oracle.scene_intersect_surfaces(sc, r)
    assert np.isfinite(t).mean() > 0.1
    assert np.all(s[np.isfinite(t)] == 0)
    # identity-placed instance under an identity transform coincides with the
    # mesh itself: the mesh (surface 0) keeps every tie
    ident = np.concatenate([np.eye(3), np.zeros((3, 1))], 1).astype(np.float32)
    sc = oracle.OracleScene([], [], [oracle.OracleMesh(vw, idx)],
                            [oracle.OracleInstance(oracle.OracleMesh(vw, idx), ident)])
    t, p, s = oracle.scene_intersect_surfaces(sc, r)
    assert np.isfinite(t).mean() > 0.1
    assert np.all(s[np.isfinite(t)] == 0)


def test_any_hit_agrees_with_closest(pkg, oracle):
    rng = np.random.default_rng(4)
    v, idx = pkg.vrt.gen_icosphere(level=3, radius=0.5, amp=0.1, freq=3.0, seed=2)
    insts = [oracle.OracleInstance(oracle.OracleMesh(v, idx), _xf(rng)) for _ in range(5)]
    sc = oracle.OracleScene([], [], [oracle.OracleMesh(v, idx)], insts)
    r = _aimed(rng, 4000, np.zeros(3), spread=1.2)
    t, p, s = oracle.scene_intersect_surfaces(sc, r)
    ta, pa, sa = oracle.scene_intersect_surfaces(sc, r, any_hit=True)
    np.testing.assert_array_equal(np.isfinite(t), np.isfinite(ta))
    assert len(set(s[np.isfinite(t)].tolist())) >= 4
