"""GPU parity of the surface path (C2): BVH closest / any hits bit-exact
against the oracle's brute force, and the mixed triangle + volume + quad
light path tracer stream-aligned with the oracle and within relMSE 1e-3 of
its mean."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(rng, n, lo=-1.5, hi=1.5):
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3:6] = d
    r[:, 7] = 1e30
    return r


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
@pytest.mark.parametrize("kind", ["soup", "icosphere"])
def test_bvh_hits_bit_exact(pkg, oracle, gpu_ctx, kind, builder):
    vrt = pkg.vrt
    rng = np.random.default_rng(6)
    if kind == "soup":
        c = rng.uniform(-1, 1, (4000, 1, 3))
        v = (c + rng.normal(0, 0.05, (4000, 3, 3))).astype(np.float32).reshape(-1, 3)
        idx = None
    else:
        v, idx = vrt.gen_icosphere(level=5, radius=0.8)
    mesh = vrt.DeviceMesh(gpu_ctx, v, idx, builder)
    print(builder, kind, "nodes", mesh.info.num_nodes, "depth", mesh.info.depth,
          "build ms", round(mesh.info.build_ms, 2))
    scene = vrt.make_scene([], [], [(mesh, (0.7, 0.7, 0.7))])
    om = oracle.OracleMesh(v, idx)
    r = _rays(rng, 20000)
    if kind == "icosphere":
        r[:5000, 0:3] *= 0.2  # rays from inside the sphere
    for any_hit in (False, True):
        t, p = vrt.intersect_surfaces(gpu_ctx, scene, r, any_hit)
        to, po = oracle.mesh_intersect(om, r, any_hit=any_hit, brute=not any_hit)
        if any_hit:
            np.testing.assert_array_equal(p >= 0, po >= 0)
        else:
            np.testing.assert_array_equal(p, po)
            np.testing.assert_array_equal(t, to)
    assert (po >= 0).mean() > 0.2


def _clustered_mesh(rng, clusters, per_cluster=24, ratio=0.5):
    """Triangle clusters along the x axis at x = ratio^i: binned SAH peels a few
    clusters off per level, so the tree is deep and its nodes have several
    inner children, which a ray along the axis hits one after another."""
    tris = []
    for i in range(clusters):
        x = ratio ** i
        s = 0.15 * x
        c = np.array([x, 0.0, 0.0]) + rng.uniform(-s, s, (per_cluster, 1, 3))
        tris.append(c + rng.normal(0.0, 0.3 * s, (per_cluster, 3, 3)))
    return np.concatenate(tris).astype(np.float32).reshape(-1, 3)


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
@pytest.mark.parametrize("clusters", [30, 60, 100])
def test_bvh_deep_clustered_tree(pkg, oracle, gpu_ctx, builder, clusters):
    """A deep tree either builds with a traversal-stack bound the kernels hold
    and then hits bit-exactly as brute force (rays along the cluster axis
    stack every sibling), or is rejected with an error -- never traversed
    with dropped subtrees."""
    vrt = pkg.vrt
    rng = np.random.default_rng(clusters)
    v = _clustered_mesh(rng, clusters)
    try:
        mesh = vrt.DeviceMesh(gpu_ctx, v, None, builder)
    except vrt.VrtError as e:
        assert "can stack" in str(e), e
        print(builder, clusters, "rejected:", e)
        return
    info = mesh.info
    print(builder, clusters, "depth", info.depth, "stack bound", info.stack_bound)
    assert info.stack_bound <= 64
    scene = vrt.make_scene([], [], [(mesh, (0.7, 0.7, 0.7))])
    n = 4000
    r = np.zeros((n, 8), np.float32)
    r[:, 7] = 1e30
    q = n // 4
    # from beyond the largest cluster towards the origin
    r[:q, 0] = rng.uniform(1.6, 2.0, q)
    r[:q, 1:3] = rng.normal(0.0, 0.02, (q, 2))
    d = np.stack([-np.ones(q), rng.normal(0, 0.01, q), rng.normal(0, 0.01, q)], 1)
    r[:q, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    # from the small end along the axis: the nearest children are the small
    # clusters, so the walk descends the deep side and stacks the larger
    # clusters at every level
    r[q:, 0] = -0.5
    r[q:, 3] = 1.0
    r[2 * q:, 1:3] = rng.normal(0.0, 1e-4, (n - 2 * q, 2))
    om = oracle.OracleMesh(v)
    t, p = vrt.intersect_surfaces(gpu_ctx, scene, r, False)
    to, po = oracle.mesh_intersect(om, r, any_hit=False, brute=True)
    np.testing.assert_array_equal(p, po)
    np.testing.assert_array_equal(t, to)
    assert (po >= 0).mean() > 0.05
    mesh.close()


def _mixed(pkg, **kw):
    return pkg.workloads.mixed_pathtrace(level=4, vol_n=32, width=64, height=40, frames=1, **kw)


def test_mixed_pathtrace_parity(pkg, oracle, gpu_ctx):
    wl = _mixed(pkg)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    ow = oracle.OracleWorkload(wl)
    for accum_id in range(2):
        gs.frame.reset()
        gs.render(accum_id)
        gpu_ctx.sync()
        g = gs.frame.color_array()
        o = (np.float32(1) / np.float32(accum_id + 1)) * ow.pathtrace_frame(accum_id)[0]
        agree = oracle.stream_agreement(g, o)
        print("mixed stream agreement", accum_id, agree)
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
        assert e <= 1e-3, (fp64, e)
    assert x.mean() > 0.01


def test_quad_light_on_volume(pkg, oracle, gpu_ctx):
    """The volume path with a quad light (MIS-weighted NEE at volume
    vertices) follows the oracle's streams."""
    wl = pkg.workloads.fbm_pathtrace(n=32, width=48, height=32, frames=1, brick=8)
    wl = dataclasses.replace(wl, lights=[dict(type="quad", position=(-0.3, 0.8, -0.3),
                                              edge1=(0.6, 0, 0), edge2=(0, 0, 0.6),
                                              color=(1, 1, 1), intensity=4.0, visible=True)])
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    ow = oracle.OracleWorkload(wl)
    gs.frame.reset()
    gs.render(0)
    gpu_ctx.sync()
    g = gs.frame.color_array()
    o = ow.pathtrace_frame(0)[0]
    agree = oracle.stream_agreement(g, o)
    print("quad light stream agreement", agree)
    assert agree >= oracle.STREAM_AGREEMENT_MIN, agree


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64])
def test_lbvh_small_and_duplicate_triangles(pkg, oracle, gpu_ctx, n):
    """Tiny meshes (a lone leaf, roots of <= 4 triangles) and repeated
    triangles (equal Morton codes, split by index) give the brute-force hits."""
    vrt = pkg.vrt
    rng = np.random.default_rng(n)
    base = rng.uniform(-0.6, 0.6, ((n + 1) // 2, 3, 3)).astype(np.float32)
    tris = np.concatenate([base, base])[:n]
    v = tris.reshape(-1, 3)
    mesh = vrt.DeviceMesh(gpu_ctx, v, None, "lbvh")
    assert mesh.info.num_triangles == n and mesh.info.num_nodes >= 1
    scene = vrt.make_scene([], [], [(mesh, (0.7, 0.7, 0.7))])
    r = _rays(rng, 4000, -1.0, 1.0)
    t, p = vrt.intersect_surfaces(gpu_ctx, scene, r, False)
    to, po = oracle.mesh_intersect(oracle.OracleMesh(v), r, brute=True)
    np.testing.assert_array_equal(p, po)
    np.testing.assert_array_equal(t, to)


def test_lbvh_rejects_bad_indices(pkg, gpu_ctx):
    vrt = pkg.vrt
    v = np.zeros((4, 3), np.float32)
    with pytest.raises(vrt.VrtError, match="out of range"):
        vrt.DeviceMesh(gpu_ctx, v, np.array([[0, 1, 2], [1, 2, 7]], np.uint32), "lbvh")


def test_mixed_scene_lbvh_same_image(pkg, gpu_ctx):
    """Hits do not depend on the tree: the GPU-built BVH renders the image of
    the host-built one."""
    wl = _mixed(pkg)
    imgs = []
    for builder in ("sah", "lbvh"):
        gs = pkg.workloads.GpuScene(gpu_ctx, wl, bvh_builder=builder)
        gs.frame.reset()
        for a in range(4):
            gs.render(a)
        gpu_ctx.sync()
        imgs.append(gs.frame.color_array())
    np.testing.assert_array_equal(imgs[0], imgs[1])
