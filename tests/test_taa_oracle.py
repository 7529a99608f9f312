"""Temporal anti-aliasing in the oracle (path tracer "taa": Pathtrace.cpp:24-25,
Pathtrace_impl.cpp:290-297, 636-637, Frame.cpp:339-381,
DeviceCopyableObjects.h:2657-2712).  A static camera gives zero motion and an
exponential blend of the frames' samples, restarted wherever the first hit's
albedo changes; a moving camera's motion vectors match the hit's pixel
positions computed from the camera's own ray basis."""
import ctypes as C
import dataclasses
import math

import numpy as np

A = 0.3


def _workload(pkg):
    wl = pkg.workloads.fbm_pathtrace(n=24, width=40, height=32, frames=1, brick=8)
    return dataclasses.replace(wl, params=dict(wl.params, max_bounce=0, ambient_samples=0))


def test_static_camera_blend(pkg, oracle):
    """Albedo is 0.8 at every hit (constant TF colour), 0 for a pixel without
    hits: frame k keeps (1 - a) history + a sample with a = 0.3 where the
    albedo matches the previous frame's, else 1."""
    wl = _workload(pkg)
    ow = oracle.OracleWorkload(wl)
    st = ow.taa_state()
    W, H = wl.width, wl.height
    hist = np.zeros((H, W, 4), np.float32)
    alb_prev = np.zeros((H, W), np.float32)
    for k in range(4):
        out, mv = ow.pathtrace_taa(k, st, A)
        assert not mv.any()
        s = ow.pathtrace_frame(k)[0]
        g = ow.gbuffer(accum_ids=(k,))
        alb = np.where(g["depth"] < 1e30, np.float32(0.8), np.float32(0))
        a = np.where(np.abs(alb - alb_prev) < 1e-2, np.float32(A), np.float32(1))[..., None]
        want = (np.float32(1) - a) * hist + a * s
        np.testing.assert_allclose(out, want, rtol=1e-6, atol=1e-7)
        hist, alb_prev = out, alb
        if k == 0:
            assert (a == 1).any() and (a == np.float32(A)).any()


def test_motion_vectors_follow_the_camera(pkg, oracle):
    """Camera moved along its right and up axes between frames: the motion
    vector (previous minus current window position of the hit; with
    maxBounce 0 the last vertex is the first hit) equals the hit's pixel
    position through the previous camera's ray basis minus the current's."""
    import test_gbuffer_oracle as G
    wl = _workload(pkg)
    cam = wl.camera
    pos = np.array(cam["position"])
    d = np.array(cam["direction"]) / np.linalg.norm(cam["direction"])
    right = np.cross(d, cam["up"])
    right /= np.linalg.norm(right)
    upv = np.cross(right, d)
    prev_pos = pos - 0.03 * right + 0.02 * upv
    prev = oracle.camera(tuple(prev_pos), cam["direction"], cam["up"], cam["fovy"], cam["aspect"])
    ow = oracle.OracleWorkload(wl)
    out, mv = ow.pathtrace_taa(0, ow.taa_state(), A, prev_cam=prev)
    g = ow.gbuffer(accum_ids=(0,))
    W, H = wl.width, wl.height
    th = math.tan(cam["fovy"] / 2)

    def window(eye, P):
        v = P - eye
        a = v @ d
        return np.array([W * ((v @ right) / (a * th * cam["aspect"]) + 1) / 2,
                         H * ((v @ upv) / (a * th) + 1) / 2])

    checked = 0
    for y in range(H):
        for x in range(W):
            if g["depth"][y, x] > 1e30:
                assert not mv[y, x].any()
                continue
            o, dv = G._primary(oracle, ow, x, y)
            P = o + float(g["depth"][y, x]) * dv
            want = window(prev_pos, P) - window(pos, P)
            np.testing.assert_allclose(mv[y, x], want, rtol=2e-3, atol=2e-3)
            checked += 1
    assert checked > 100
    assert np.abs(mv).max() > 0.5  # a real shift of about a pixel or more
