"""Gradient shading in the oracle (renderer "gradientShading", Renderer.cpp:50):
the ray caster's matte shading along the central-difference gradient
(VolumeIntegration.h:420-459, DeviceCopyableObjects.h:582-597) and the path
tracer's gradient normal and matte NEE (Pathtrace_impl.cpp:204-212,
311-328).  On a field affine in x the gradient normal is +-x exactly, so each
pixel's shading factor is a closed form."""
import dataclasses
import math

import numpy as np
import pytest

L_DIR = np.array([0.8, -0.3, -0.5])  # direction the light travels


def _affine_workload(pkg, renderer="raycast", n=24, width=40, height=32, **params):
    """value = normalised x over the volume; TF rgb a ramp, alpha 0.5 only for
    values in [0.3, 0.7] (well inside, where central differences are exact)."""
    wls = pkg.workloads
    x = np.linspace(0.0, 1.0, n, dtype=np.float32)
    vox = np.broadcast_to(x[None, None, :], (n, n, n)).copy()
    tf = np.zeros((64, 4), np.float32)
    u = np.linspace(0.0, 1.0, 64)
    tf[:, 0], tf[:, 1], tf[:, 2] = u, 0.5, 1.0 - u
    tf[(u >= 0.3) & (u <= 0.7), 3] = 0.5
    pos = (0.0, 0.3, 2.0)  # rays on both sides of x = 0: both normal signs
    prm = dict(max_bounce=0, ambient_samples=0, ambient_radiance=0.3,
               ambient_color=(1.0, 0.9, 0.8), volume_sampling_rate=1.0)
    prm.update(params)
    return wls.Workload(
        name="affine x", voxels=vox, origin=(-0.5,) * 3, spacing=(1.0 / (n - 1),) * 3, tf=tf,
        value_range=(0.0, 1.0), unit_distance=0.05, grid_dims=None,
        camera=dict(position=pos, direction=wls.look_at(pos), up=(0, 1, 0),
                    fovy=math.radians(45.0), aspect=width / height),
        lights=[dict(type="directional", direction=tuple(L_DIR), color=(1, 1, 1), intensity=2.0)],
        params=prm, width=width, height=height, renderer=renderer)


def test_raycast_gradient_factor_closed_form(pkg, oracle):
    """Every contributing sample has normal sign(-dir.x) * x, so the shaded
    pixel is the unshaded one times E/pi * max(0, n.L) + ambient."""
    import ctypes as C
    wl = _affine_workload(pkg)
    plain = oracle.OracleWorkload(wl).raycast()
    ow = oracle.OracleWorkload(dataclasses.replace(wl, params=dict(wl.params,
                                                                   gradient_shading=True)))
    shaded = ow.raycast()
    Ld = -L_DIR / np.linalg.norm(L_DIR)
    amb = np.array([1.0, 0.9, 0.8]) * 0.3
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    checked = signs = 0
    for y in range(wl.height):
        for x in range(wl.width):
            if plain[y, x, :3].max() < 1e-4:
                continue
            oracle.lib().orc_camera_ray(C.byref(ow.cam), x, y, wl.width, wl.height,
                                        o.ctypes.data, d.ctypes.data)
            nx = 1.0 if -d[0] >= 0 else -1.0
            k = 2.0 / math.pi * max(0.0, nx * Ld[0]) + amb
            np.testing.assert_allclose(shaded[y, x, :3], plain[y, x, :3] * k, rtol=2e-5,
                                       atol=1e-7)
            assert shaded[y, x, 3] == plain[y, x, 3]
            checked += 1
            signs += nx > 0
    assert checked > 200 and 0 < signs < checked


def test_pathtrace_gradient_nee_ratio(pkg, oracle):
    """Path tracer, first vertex only (maxBounce 0, no AO): with gradient
    shading the NEE is albedo/pi * E * max(0, sn.L) * numLights instead of
    albedo * E -- per pixel the ratio max(0, sn.L)/pi, sn = sign(-dir.x) x;
    the g-buffer normal is that sn."""
    import ctypes as C
    import test_gbuffer_oracle as G
    wl = _affine_workload(pkg, "pathtrace", ambient_radiance=0.0)
    plain = oracle.OracleWorkload(wl).pathtrace_frame(0)[0]
    gw = dataclasses.replace(wl, params=dict(wl.params, gradient_shading=True))
    ow = oracle.OracleWorkload(gw)
    grad = ow.pathtrace_frame(0)[0]
    g = ow.gbuffer()
    Ld = -L_DIR / np.linalg.norm(L_DIR)
    checked = 0
    for y in range(wl.height):
        for x in range(wl.width):
            if plain[y, x, 0] <= 0 or g["depth"][y, x] > 1e30:
                continue
            _, d = G._primary(oracle, ow, x, y)
            nx = 1.0 if -d[0] >= 0 else -1.0
            np.testing.assert_allclose(g["normal"][y, x], [nx, 0, 0], rtol=0, atol=1e-6)
            want = max(0.0, nx * Ld[0]) / math.pi
            np.testing.assert_allclose(grad[y, x, :3] / plain[y, x, :3], want, rtol=2e-5)
            checked += 1
    assert checked > 100


def test_constant_field_has_no_gradient(pkg, oracle):
    """A flat field has a zero gradient: no normal, so without ambient
    occlusion gradient shading changes nothing (the reference would normalise
    the zero vector to NaN).  With ambient occlusion the random AO normal is
    drawn instead and, being a normal, selects the matte NEE
    (Pathtrace_impl.cpp:213-215, 313)."""
    wl = _affine_workload(pkg, "pathtrace", max_bounce=2, ambient_samples=0)
    # value 1 everywhere: every trilinear lerp of (1, 1) is exactly 1, since
    # fl(1 - w) + w rounds to 1 -- the central differences are exactly 0
    tf = wl.tf.copy()
    tf[:, 3] = 0.5
    wl = dataclasses.replace(wl, voxels=np.ones_like(wl.voxels), tf=tf)
    a = oracle.OracleWorkload(wl).pathtrace_frame(0)[0]
    b = oracle.OracleWorkload(dataclasses.replace(
        wl, params=dict(wl.params, gradient_shading=True))).pathtrace_frame(0)[0]
    assert np.abs(a).max() > 0
    np.testing.assert_array_equal(a, b)
