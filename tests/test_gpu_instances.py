"""Instanced meshes on the GPU (vrt_tls: the instance-level BVH and
intersect(Ray, Instance), DeviceCopyableObjects.h:1772-1824) against the
oracle's instances, given the same transforms:

  hits        closest hits bit-exact (t, triangle, which surface) over a plain
              mesh, rotated / non-uniformly scaled instances, coincident
              instances (the lower index wins) and an identity instance
              coinciding with the plain mesh (the mesh wins); any-hit masks
              equal; a TLS of 300 instances the same way
  ray caster  the instanced scene's image within 1e-3 of the oracle's
  tracer      stream-aligned with the oracle and within relMSE 1e-3 of its
              mean; g-buffer ids / normals of instanced hits equal
"""
import dataclasses
import math

import numpy as np
import pytest

import test_gpu_gbuffer as GB

pytestmark = pytest.mark.gpu


def _xf(rng, lo=0.3, hi=1.2, spread=1.0):
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    ang = rng.uniform(0, 2 * math.pi)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)
    L = R @ np.diag(rng.uniform(lo, hi, 3))
    return np.concatenate([L, rng.uniform(-spread, spread, (3, 1))], 1).astype(np.float32)


def _aimed(rng, n, spread=1.5, dist=4.0):
    o = rng.normal(size=(n, 3))
    o = dist * o / np.linalg.norm(o, axis=1, keepdims=True)
    d = rng.uniform(-spread, spread, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3:6] = d
    r[:, 7] = 1e30
    return r


def _scene_pair(pkg, oracle, ctx, v, idx, xfs, plain=True):
    vrt = pkg.vrt
    mesh = vrt.DeviceMesh(ctx, v, idx)
    tls = vrt.DeviceTls(ctx, [(mesh, x, (0.5, 0.6, 0.7), 7, 100 + k) for k, x in enumerate(xfs)])
    scene = vrt.make_scene([], [], [(mesh, (0.7, 0.7, 0.7))] if plain else [], tls)
    osc = oracle.OracleScene([], [], [oracle.OracleMesh(v, idx)] if plain else [],
                             [oracle.OracleInstance(oracle.OracleMesh(v, idx), x) for x in xfs])
    return mesh, tls, scene, osc


def _check_hits(pkg, oracle, ctx, scene, osc, r):
    vrt = pkg.vrt
    t, p, s = vrt.intersect_surfaces(ctx, scene, r, False, which=True)
    to, po, so = oracle.scene_intersect_surfaces(osc, r)
    np.testing.assert_array_equal(s, so)
    np.testing.assert_array_equal(p, po)
    np.testing.assert_array_equal(t, to)
    ta, pa = vrt.intersect_surfaces(ctx, scene, r, True)
    np.testing.assert_array_equal(pa >= 0, po >= 0)
    return s


def test_instance_hits_bit_exact(pkg, oracle, gpu_ctx):
    vrt = pkg.vrt
    rng = np.random.default_rng(11)
    v, idx = vrt.gen_icosphere(level=3, radius=0.6, amp=0.1, freq=3.0, seed=5)
    ident = np.concatenate([np.eye(3), np.zeros((3, 1))], 1).astype(np.float32)
    xfs = [_xf(rng) for _ in range(6)]
    xfs += [xfs[2], xfs[2], ident]  # coincident instances; one on the plain mesh
    mesh, tls, scene, osc = _scene_pair(pkg, oracle, gpu_ctx, v, idx, xfs)
    print("tls", tls.info.num_instances, "nodes", tls.info.num_nodes,
          "stack bound", tls.info.stack_bound)
    s = _check_hits(pkg, oracle, gpu_ctx, scene, osc, _aimed(rng, 30000))
    hit = s >= 0
    assert hit.mean() > 0.3
    seen = set(s[hit].tolist())
    assert {0, 1, 2, 3, 4, 5, 6}.issubset(seen)
    assert 1 + 6 not in seen and 1 + 7 not in seen and 1 + 8 not in seen  # ties kept
    tls.close()
    mesh.close()


def test_many_instances(pkg, oracle, gpu_ctx):
    vrt = pkg.vrt
    rng = np.random.default_rng(12)
    v, idx = vrt.gen_icosphere(level=2, radius=1.0, amp=0.1, freq=3.0, seed=6)
    xfs = [_xf(rng, 0.05, 0.15, spread=1.5) for _ in range(300)]
    mesh, tls, scene, osc = _scene_pair(pkg, oracle, gpu_ctx, v, idx, xfs, plain=False)
    print("tls", tls.info.num_instances, "nodes", tls.info.num_nodes,
          "stack bound", tls.info.stack_bound)
    assert tls.info.num_nodes > 20
    s = _check_hits(pkg, oracle, gpu_ctx, scene, osc, _aimed(rng, 30000))
    assert (s >= 0).mean() > 0.2 and len(set(s[s >= 0].tolist())) > 100
    tls.close()
    mesh.close()


def test_tls_rejects_bad_input(pkg, gpu_ctx):
    vrt = pkg.vrt
    v, idx = vrt.gen_icosphere(level=1, radius=1.0, amp=0.0, freq=1.0, seed=1)
    mesh = vrt.DeviceMesh(gpu_ctx, v, idx)
    sing = np.zeros((3, 4), np.float32)
    with pytest.raises(vrt.VrtError, match="singular"):
        vrt.DeviceTls(gpu_ctx, [(mesh, sing, (1, 1, 1))])
    with pytest.raises(vrt.VrtError):
        vrt.DeviceTls(gpu_ctx, [])
    mesh.close()


def _wl(pkg, renderer, **kw):
    wl = pkg.workloads.instanced_pathtrace(width=96, height=72, renderer=renderer, **kw)
    return wl


def test_instanced_raycast_parity(pkg, oracle, gpu_ctx):
    wl = _wl(pkg, "raycast")
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    gs.render()
    gpu_ctx.sync()
    g = gs.frame.color_array()
    o = oracle.OracleWorkload(wl).raycast()
    err = np.abs(g - o)
    print("instanced raycast max-abs", err.max())
    assert err.max() <= 1e-3
    # the instances are in the picture, in their own colours
    assert len(np.unique(np.round(g[..., :3], 2).reshape(-1, 3), axis=0)) > 50


def test_instanced_pathtrace_parity(pkg, oracle, gpu_ctx):
    wl = _wl(pkg, "pathtrace")
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    ow = oracle.OracleWorkload(wl)
    for accum_id in range(2):
        gs.frame.reset()
        gs.render(accum_id)
        gpu_ctx.sync()
        g = gs.frame.color_array()
        o = (np.float32(1) / np.float32(accum_id + 1)) * ow.pathtrace_frame(accum_id)[0]
        agree = oracle.stream_agreement(g, o)
        print("instanced stream agreement", accum_id, agree)
        assert agree >= oracle.STREAM_AGREEMENT_MIN, agree
    frames = 32
    gs.frame.reset()
    for a in range(frames):
        gs.render(a)
    gpu_ctx.sync()
    x = np.asarray(gs.frame.accum_array(), np.float64)[..., :3]
    r = np.asarray(ow.pathtrace_accumulate(frames), np.float64)[..., :3]
    e = float(np.mean((x - r) ** 2 / (r * r + 1e-2)))
    print("instanced relMSE", e)
    assert e <= 1e-3
    assert x.mean() > 0.01


def test_instanced_gbuffer(pkg, oracle, gpu_ctx):
    """First-hit ids of instanced meshes: object id 7 and instance ids
    100 + k, normals through the instances' normal transforms."""
    wl = _wl(pkg, "pathtrace")
    g = GB._check(pkg, oracle, gpu_ctx, wl, frames=2)
    ids = set(np.unique(g["instance_id"]).tolist())
    assert len(ids & set(range(100, 106))) >= 4
