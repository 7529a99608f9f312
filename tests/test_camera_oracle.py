"""Pins of the oracle's perspective camera (camera/Perspective.cpp:14-44,
DeviceCopyableObjects.h:2583-2600; visionaray's thin_lens_camera itself is
not vendored) from the geometry of a pinhole and of a thin lens:

* the pinhole ray of pixel (x, y) starts at the eye and passes through the
  point that the textbook look-at + perspective projection (gluLookAt,
  gluPerspective: fovy the full vertical angle, aspect = width / height)
  maps to normalised device coordinates 2 s - 1, with s the pixel centre
  ((x + .5) / W, (y + .5) / H) carried into imageRegion (row 0 at the
  bottom);
* a thin-lens ray starts on the lens disk (radius apertureRadius, in the
  plane through the eye perpendicular to the view direction) and passes
  through the pinhole ray's point on the plane in focus, focusDistance along
  the view direction; the concentric map keeps the lens samples uniform
  over the disk (E|p|^2 = R^2 / 2, a quarter inside R / 2, centred).
"""
import ctypes as C
import math

import numpy as np
import pytest


def _normalize(v):
    return v / np.linalg.norm(v)


def _look_at(eye, center, up):
    """gluLookAt: world -> eye space, the viewer looking down -z."""
    f = _normalize(center - eye)
    s = _normalize(np.cross(f, up))
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def _perspective(fovy, aspect, zn=0.001, zf=1000.0):
    """gluPerspective."""
    f = 1.0 / math.tan(fovy / 2.0)
    return np.array([[f / aspect, 0, 0, 0], [0, f, 0, 0],
                     [0, 0, (zf + zn) / (zn - zf), 2 * zf * zn / (zn - zf)], [0, 0, -1, 0]])


def _cameras():
    rng = np.random.default_rng(3)
    out = []
    for k in range(6):
        pos = rng.uniform(-3, 3, 3)
        d = _normalize(rng.normal(size=3))
        up = _normalize(rng.normal(size=3))
        if abs(np.dot(d, up)) > 0.9:
            up = _normalize(np.cross(d, rng.normal(size=3)))
        fovy = float(rng.uniform(0.3, 1.8))
        aspect = float(rng.uniform(0.5, 2.0))
        region = (0, 0, 1, 1) if k % 2 == 0 else (0.2, 0.1, 0.7, 0.9)
        # the camera holds float32 parameters: compare with what it holds
        pos, d, up = (v.astype(np.float32).astype(np.float64) for v in (pos, d, up))
        out.append((pos, d, up, fovy, aspect, region))
    return out


def _oracle_camera(oracle, pos, d, up, fovy, aspect, region, radius=0.0, focus=1.0):
    return oracle.camera(tuple(pos), tuple(d), tuple(up), fovy, aspect, region=region,
                         aperture_radius=radius, focus_distance=focus)


@pytest.mark.parametrize("k", range(6))
def test_pinhole_matches_look_at_perspective(oracle, k):
    pos, d, up, fovy, aspect, region = _cameras()[k]
    cam = _oracle_camera(oracle, pos, d, up, fovy, aspect, region)
    mvp = _perspective(fovy, aspect) @ _look_at(pos, pos + d, up)
    W, H = 9, 7
    o, r = np.empty(3, np.float32), np.empty(3, np.float32)
    for y in range(H):
        for x in range(W):
            oracle.lib().orc_camera_ray(C.byref(cam), float(x), float(y), W, H,
                                        o.ctypes.data, r.ctypes.data)
            np.testing.assert_allclose(o, pos, atol=1e-6)
            assert abs(np.linalg.norm(r) - 1.0) < 1e-6
            clip = mvp @ np.append(o.astype(np.float64) + 2.5 * r, 1.0)
            assert clip[3] > 0  # in front of the eye
            ndc = clip[:2] / clip[3]
            sx = (x + 0.5) / W
            sy = (y + 0.5) / H
            sx = (1 - sx) * region[0] + sx * region[2]
            sy = (1 - sy) * region[1] + sy * region[3]
            np.testing.assert_allclose(ndc, [2 * sx - 1, 2 * sy - 1], atol=2e-5)


def _lens_ray(oracle, cam, x, y, W, H, u1, u2):
    o, r = np.empty(3, np.float64), np.empty(3, np.float64)
    oracle.lib().orc_camera_lens_ray(C.byref(cam), float(x), float(y), W, H, float(u1),
                                     float(u2), o.ctypes.data, r.ctypes.data)
    return o, r


@pytest.mark.parametrize("k", [0, 3])
def test_thin_lens_geometry(oracle, k):
    pos, d, up, fovy, aspect, region = _cameras()[k]
    d = _normalize(d)  # the view direction the camera derives from what it holds
    radius, focus = float(np.float32(0.15)), float(np.float32(2.7))  # as the camera holds them
    cam = _oracle_camera(oracle, pos, d, up, fovy, aspect, region, radius, focus)
    W, H = 8, 6
    n = 24
    us = (np.arange(n) + 0.5) / n
    for (x, y) in [(0, 0), (3, 2), (7, 5)]:
        # the lens centre (u = 1/2) is the pinhole ray
        o0, r0 = _lens_ray(oracle, cam, x, y, W, H, 0.5, 0.5)
        p0, q0 = np.empty(3, np.float32), np.empty(3, np.float32)
        oracle.lib().orc_camera_ray(C.byref(cam), float(x), float(y), W, H, p0.ctypes.data,
                                    q0.ctypes.data)
        np.testing.assert_allclose(o0, pos, atol=1e-6)
        np.testing.assert_allclose(r0, q0, atol=1e-6)
        # the pinhole ray's point on the plane in focus
        fp = o0 + r0 * (focus / np.dot(r0, d))
        assert abs(np.dot(fp - pos, d) - focus) < 1e-9
        lens = []
        for u1 in us:
            for u2 in us:
                o, r = _lens_ray(oracle, cam, x, y, W, H, u1, u2)
                off = o - pos
                assert abs(np.dot(off, d)) < 1e-9          # lens plane
                assert np.linalg.norm(off) <= radius * (1 + 1e-9)
                assert abs(np.linalg.norm(r) - 1.0) < 1e-12
                t = np.dot(fp - o, r)
                assert np.linalg.norm(o + t * r - fp) < 1e-9  # through the focus point
                lens.append(off)
        lens = np.array(lens) / radius
        r2 = np.sum(lens ** 2, axis=1)
        assert abs(r2.mean() - 0.5) < 0.01
        assert abs(np.mean(r2 < 0.25) - 0.25) < 0.02
        assert np.linalg.norm(lens.mean(axis=0)) < 0.01


def test_zero_aperture_is_the_pinhole(oracle):
    """Radius 0 leaves the pinhole ray (the renderers then draw no lens
    sample), whatever the focus distance."""
    pos, d, up, fovy, aspect, region = _cameras()[1]
    a = _oracle_camera(oracle, pos, d, up, fovy, aspect, region, 0.0, 5.0)
    for u1, u2 in [(0.1, 0.9), (0.7, 0.3)]:
        o, r = _lens_ray(oracle, a, 2, 3, 8, 6, u1, u2)
        p, q = np.empty(3, np.float32), np.empty(3, np.float32)
        oracle.lib().orc_camera_ray(C.byref(a), 2.0, 3.0, 8, 6, p.ctypes.data, q.ctypes.data)
        np.testing.assert_allclose(o, p, atol=1e-6)
        np.testing.assert_allclose(r, q, atol=1e-6)
