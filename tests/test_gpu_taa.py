"""Temporal anti-aliasing on the GPU against the oracle: frame-by-frame
resolved images and motion vectors, static and moving camera
(Frame.cpp:339-381, DeviceCopyableObjects.h:2657-2712)."""
import dataclasses

import numpy as np
import pytest

import test_taa_oracle as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("moving", [False, True])
def test_taa_frames_match_oracle(pkg, oracle, gpu_ctx, moving):
    vrt = pkg.vrt
    wl = T._workload(pkg)
    wl = dataclasses.replace(wl, params=dict(wl.params, max_bounce=3, ambient_samples=1,
                                             taa=True, taa_alpha=0.3))
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, taa=True)
    ow = oracle.OracleWorkload(wl)
    st = ow.taa_state()
    cam = wl.camera
    pos = np.array(cam["position"])
    prev = None
    for k in range(4):
        p = tuple(pos + (0.01 * k * np.array([1.0, 0.5, 0.0]) if moving else 0.0))
        c_args = dict(position=p, direction=cam["direction"], up=cam["up"], fovy=cam["fovy"],
                      aspect=cam["aspect"])
        gs.camera = vrt.camera(**c_args)
        gs.params = vrt.params(**wl.params, prev_camera=vrt.camera(**prev) if prev else None)
        gs.render(k)
        gpu_ctx.sync()
        img = gs.frame.color_array()
        mv = gs.frame.taa["motion_vec"].download(np.float32, (wl.height, wl.width, 4))[..., :2]
        ocam = oracle.camera(**c_args)
        oprev = oracle.camera(**prev) if prev else None
        out, omv = ow.pathtrace_taa(k, st, 0.3, cam=ocam, prev_cam=oprev)
        agree = oracle.stream_agreement(img, out)
        mv_ok = np.isclose(mv, omv, rtol=1e-4, atol=1e-4).all(-1).mean()
        print("taa frame", k, "moving", moving, "agreement", agree, "motion", mv_ok)
        assert agree >= oracle.STREAM_AGREEMENT_MIN
        assert mv_ok >= oracle.STREAM_AGREEMENT_MIN
        prev = c_args
    if moving:
        assert np.abs(mv).max() > 0.1


def test_taa_needs_buffers(pkg, gpu_ctx):
    wl = T._workload(pkg)
    wl = dataclasses.replace(wl, params=dict(wl.params, taa=True))
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)  # no taa buffers
    with pytest.raises(pkg.vrt.VrtError, match="taa"):
        gs.render(0)
