"""HDRI environment light on the GPU against the oracle: path tracing with
NEE, light hits and MIS (volume and mixed surface scenes), and the ray
caster's light backdrop (HDRI and quad)."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _env(seed=5, w=32, h=16, up=(0, 1, 0), direction=(0.3, 0, 1), scale=1.2, visible=True):
    rng = np.random.default_rng(seed)
    m = rng.uniform(0.05, 0.6, (h, w, 3)).astype(np.float32)
    m[h // 4, w // 3] = (40.0, 35.0, 30.0)  # a sun-like cell
    return dict(type="hdri", radiance=m, up=up, direction=direction, scale=scale, visible=visible)


def _agree(pkg, oracle, gpu_ctx, wl, frames=1):
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    ow = oracle.OracleWorkload(wl)
    gs.frame.reset()
    worst = 1.0
    for a in range(frames):
        gs.render(a)
        gpu_ctx.sync()
        img = gs.frame.color_array()
        ref = ow.pathtrace_frame(a)[0] if a == 0 else None
        if ref is not None:
            worst = min(worst, oracle.stream_agreement(img, ref))
    return worst, gs.frame.color_array()


@pytest.mark.parametrize("visible", [True, False])
def test_pathtrace_volume_under_hdri(pkg, oracle, gpu_ctx, visible):
    wl = pkg.workloads.fbm_pathtrace(n=32, width=64, height=48, frames=1, brick=8)
    wl = dataclasses.replace(wl, lights=[_env(visible=visible)])
    agree, img = _agree(pkg, oracle, gpu_ctx, wl)
    print("hdri volume stream agreement", agree, "visible", visible)
    assert agree >= oracle.STREAM_AGREEMENT_MIN
    # the map shows behind the volume only when visible to camera rays
    bg = np.array(wl.params.get("background", (0, 0, 0, 1)), np.float32)
    assert (np.abs(img[0, 0] - bg).max() > 1e-3) == visible


def test_pathtrace_mixed_scene_hdri_and_quad(pkg, oracle, gpu_ctx):
    wl = pkg.workloads.mixed_pathtrace(level=3, vol_n=24, width=48, height=32, frames=1)
    wl = dataclasses.replace(wl, lights=list(wl.lights) + [_env(up=(0, 0, 1), direction=(1, 0, 0))])
    agree, _ = _agree(pkg, oracle, gpu_ctx, wl)
    print("hdri mixed stream agreement", agree)
    assert agree >= oracle.STREAM_AGREEMENT_MIN


@pytest.mark.parametrize("backdrop", ["hdri", "quad"])
def test_raycast_light_backdrop(pkg, oracle, gpu_ctx, backdrop):
    wl = pkg.workloads.blobs_raycast(n=32, width=96, height=64)
    if backdrop == "hdri":
        light = _env()
    else:
        light = dict(type="quad", position=(-4.0, -3.0, -1.0), edge1=(8.0, 0, 0),
                     edge2=(0, 6.0, 0), intensity=2.0, color=(0.9, 0.8, 0.7), visible=True)
    wl = dataclasses.replace(wl, lights=[light])
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    gs.render()
    gpu_ctx.sync()
    g = gs.frame.color_array()
    o = oracle.OracleWorkload(wl).raycast()
    err = np.abs(g - o).max()
    print(backdrop, "backdrop raycast max-abs", err)
    assert err <= 1e-3
    assert (np.abs(g[..., :3]).max(-1) > 0.01).mean() > 0.9  # the light, not the background
