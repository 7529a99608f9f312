"""Surfaces in the oracle (C2): triangle intersection, the oracle's BVH, quad
light sampling and the matte shading path, each against closed forms or
brute force (Pathtrace_impl.cpp:93-487, renderer/common.h:1033-1048,
1371-1466, DeviceCopyableObjects.h:2187-2273)."""
import math

import numpy as np


def _rays(o, d, tmin=0.0, tmax=1e30):
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    n = max(len(o), len(d))
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3:6] = d
    r[:, 6] = tmin
    r[:, 7] = tmax
    return r


def test_triangle_closed_form(oracle):
    """One triangle in z = 0 seen along -z: hit exactly where x, y >= 0 and
    x + y <= 1 (edges included) at t = 2; parallel rays and hits outside
    [tmin, tmax] miss."""
    mesh = oracle.OracleMesh([[0, 0, 0], [1, 0, 0], [0, 1, 0]], color=(1, 1, 1))
    g = np.arange(-0.25, 1.3, 0.125, dtype=np.float32)
    xy = np.array([(x, y) for x in g for y in g], np.float32)
    o = np.column_stack([xy, np.full(len(xy), 2.0, np.float32)])
    t, p = oracle.mesh_intersect(mesh, _rays(o, [[0, 0, -1]] * len(o)), brute=True)
    inside = (xy[:, 0] >= 0) & (xy[:, 1] >= 0) & (xy[:, 0] + xy[:, 1] <= 1)
    np.testing.assert_array_equal(p >= 0, inside)
    np.testing.assert_array_equal(t[inside], np.float32(2))
    t, p = oracle.mesh_intersect(mesh, _rays([[0.2, 0.2, 1]], [[1, 0, 0]]), brute=True)
    assert p[0] == -1
    t, p = oracle.mesh_intersect(mesh, _rays([[0.2, 0.2, 2]], [[0, 0, -1]], 0.0, 1.5), brute=True)
    assert p[0] == -1
    t, p = oracle.mesh_intersect(mesh, _rays([[0.2, 0.2, 2]], [[0, 0, -1]], 2.5, 9.0), brute=True)
    assert p[0] == -1


def test_bvh_matches_brute_force(oracle):
    rng = np.random.default_rng(4)
    n = 3000
    c = rng.uniform(-1, 1, (n, 1, 3))
    tris = (c + rng.normal(0, 0.06, (n, 3, 3))).astype(np.float32).reshape(-1, 3)
    mesh = oracle.OracleMesh(tris)
    o = rng.uniform(-1.5, 1.5, (4000, 3))
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = _rays(o, d)
    tb, pb = oracle.mesh_intersect(mesh, r, brute=True)
    t, p = oracle.mesh_intersect(mesh, r)
    np.testing.assert_array_equal(p, pb)
    np.testing.assert_array_equal(t, tb)
    assert (p >= 0).mean() > 0.2
    ta, pa = oracle.mesh_intersect(mesh, r, any_hit=True)
    np.testing.assert_array_equal(pa >= 0, pb >= 0)


def test_quad_light_samples(oracle):
    """Samples are uniform over the quad (edge coordinates: mean 1/2,
    variance 1/12) and pdf * area * |cos theta_light| = 1."""
    L = oracle.light_quad((-1, 1, 0), (2, 0, 0), (0, 0, 1), intensity=3.0)
    ref = np.array([0.3, 0.0, 0.2], np.float32)
    pos, pdf = oracle.quad_light_sample(L, ref, 200000, seed=7)
    s = (pos[:, 0] + 1) / 2
    t = pos[:, 2]
    assert np.allclose(pos[:, 1], 1.0)
    assert s.min() >= -1e-6 and s.max() <= 1 + 1e-6 and t.min() >= -1e-6 and t.max() <= 1 + 1e-6
    for u in (s, t):
        assert abs(u.mean() - 0.5) < 4e-3
        assert abs(u.var() - 1 / 12) < 2e-3
    d = pos - ref
    cosl = np.abs(d[:, 1]) / np.linalg.norm(d, axis=1)
    np.testing.assert_allclose(pdf * 2.0 * cosl, 1.0, rtol=2e-6)


def _plane_scene(oracle, lights, params, color=(0.6, 0.5, 0.4), W=8, H=6):
    v = np.array([[-5, 0, -5], [5, 0, -5], [5, 0, 5], [-5, 0, 5]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
    mesh = oracle.OracleMesh(v, idx, color)
    scene = oracle.OracleScene([], lights, [mesh])
    cam = oracle.camera((0.1, 1.0, 0.3), (0.0, -1.0, -0.1), (0, 0, -1), math.radians(30), W / H)
    prm = oracle.params(**params)
    out = np.zeros((H, W, 4), np.float32)
    oracle.lib().orc_pathtrace_frame(oracle.C.byref(scene.s), oracle.C.byref(cam),
                                     oracle.C.byref(prm), W, H, 0, 0, 0, 1, 0, 0,
                                     out.ctypes.data, 1, None)
    return out


def test_matte_closed_forms(oracle):
    """A matte plane: lit only by the ambient term it returns color * ambient
    (no occluder, AO visibility 1); lit by a directional light of irradiance
    E at angle theta it returns color / pi * E * cos(theta) (delta light: no
    MIS, one bounce)."""
    c = np.array([0.6, 0.5, 0.4])
    out = _plane_scene(oracle, [], dict(ambient_color=(1, 0.5, 0.25), ambient_radiance=0.8,
                                        ambient_samples=1, max_bounce=0))
    np.testing.assert_allclose(out[..., :3], np.broadcast_to(c * [1, 0.5, 0.25] * 0.8, out[..., :3].shape),
                               rtol=1e-6)
    theta = math.radians(35)
    d = (math.sin(theta), -math.cos(theta), 0.0)  # light travel direction
    L = oracle.light_directional(d, 2.5, (1, 1, 1))
    out = _plane_scene(oracle, [L], dict(ambient_radiance=0.0, ambient_samples=0, max_bounce=0))
    exp = c / math.pi * 2.5 * math.cos(theta)
    np.testing.assert_allclose(out[..., :3], np.broadcast_to(exp, out[..., :3].shape), rtol=2e-6)


def _rect_form_factor(x1, x2, z1, z2, h):
    """Form factor from a differential area at the origin (normal +y) to the
    rectangle [x1, x2] x [z1, z2] in the plane y = h: the corner-rectangle
    formula (Howell's catalogue, differential element to a parallel
    rectangle) combined by inclusion-exclusion; it is odd in each side."""
    def corner(X, Z):
        a, b = math.sqrt(1 + X * X), math.sqrt(1 + Z * Z)
        return (X / a * math.atan(Z / a) + Z / b * math.atan(X / b)) / (2 * math.pi)
    return (corner(x2 / h, z2 / h) - corner(x1 / h, z2 / h) - corner(x2 / h, z1 / h)
            + corner(x1 / h, z1 / h))


def test_ambient_occlusion_form_factor(oracle):
    """AO rays leave a matte point in a cosine-weighted hemisphere, so the
    chance that one hits a parallel rectangle above it is that rectangle's
    form factor F (0.162 here; a uniform hemisphere would give 0.126).  The
    mean AO-lit value is color * ambient * (1 - F) within 4 binomial sigma."""
    c = np.array([0.6, 0.5, 0.4])
    h, x1, x2, z1, z2 = 0.5, 0.25, 1.25, -0.5, 0.5
    plane = oracle.OracleMesh(np.array([[-5, 0, -5], [5, 0, -5], [5, 0, 5], [-5, 0, 5]], np.float32),
                              np.array([[0, 2, 1], [0, 3, 2]], np.uint32), c)
    occ = oracle.OracleMesh(np.array([[x1, h, z1], [x2, h, z1], [x2, h, z2], [x1, h, z2]], np.float32),
                            np.array([[0, 1, 2], [0, 2, 3]], np.uint32), (0.7, 0.7, 0.7))
    scene = oracle.OracleScene([], [], [plane, occ])
    # a near-zero field of view: every camera ray lands within 1e-3 of the origin
    cam = oracle.camera((0, 1, 0), (0, -1, 0), (0, 0, -1), 1e-3, 1.0)
    prm = oracle.params(ambient_radiance=1.0, ambient_samples=1, max_bounce=0)
    W = H = 32
    frames = 40
    out = np.zeros((H, W, 4), np.float64)
    oracle.lib().orc_pathtrace_accumulate(oracle.C.byref(scene.s), oracle.C.byref(cam),
                                          oracle.C.byref(prm), W, H, 0, frames, 0, 1,
                                          out.ctypes.data, 4)
    F = _rect_form_factor(x1, x2, z1, z2, h)
    n = W * H * frames
    vis = out[..., :3].mean(axis=(0, 1)) / c
    sigma = math.sqrt(F * (1 - F) / n)
    print("AO visibility", vis, "expected", 1 - F)
    assert np.all(np.abs(vis - (1 - F)) <= 4 * sigma), (vis, 1 - F, sigma)


def test_visible_quad_light_closed_form(oracle):
    """Camera rays that hit a visible quad light return its intensity (no MIS
    at the first vertex); an invisible one leaves the background."""
    for visible, exp in ((True, 3.0), (False, 0.0)):
        L = oracle.light_quad((-2, -2, -1), (4, 0, 0), (0, 4, 0), intensity=3.0, visible=visible)
        scene = oracle.OracleScene([], [L], [oracle.OracleMesh([[9, 9, 9], [9, 9.1, 9], [9.1, 9, 9]])])
        cam = oracle.camera((0, 0, 1), (0, 0, -1), (0, 1, 0), math.radians(20), 1.0)
        prm = oracle.params()
        out = np.zeros((4, 4, 4), np.float32)
        oracle.lib().orc_pathtrace_frame(oracle.C.byref(scene.s), oracle.C.byref(cam),
                                         oracle.C.byref(prm), 4, 4, 0, 0, 0, 1, 0, 0,
                                         out.ctypes.data, 1, None)
        np.testing.assert_array_equal(out[..., :3], np.float32(exp))


def test_mixed_scene_renders(pkg, oracle):
    wl = pkg.workloads.mixed_pathtrace(level=3, vol_n=24, width=24, height=16, frames=1)
    acc = oracle.OracleWorkload(wl).pathtrace_accumulate(16)
    assert np.all(np.isfinite(acc))
    assert acc[..., :3].mean() > 0.01
