"""Gradient shading on the GPU against the oracle: the ray caster's matte
shading along the central-difference gradient (VolumeIntegration.h:420-459)
for structured, NanoVDB and tetrahedral fields and directional / quad
lights, and the path tracer's gradient normal and matte NEE
(Pathtrace_impl.cpp:204-212, 311-328)."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SUN = dict(type="directional", direction=(0.5, -0.7, -0.4), color=(1.0, 0.95, 0.9),
           intensity=2.5)


def _shaded(wl, lights=(SUN,), **params):
    return dataclasses.replace(wl, lights=list(lights),
                               params={**wl.params, "gradient_shading": True, **params})


def _raycast_pair(pkg, oracle, gpu_ctx, wl):
    gs = pkg.workloads.GpuScene(gpu_ctx, wl)
    gs.render()
    gpu_ctx.sync()
    return gs.frame.color_array(), oracle.OracleWorkload(wl).raycast()


@pytest.mark.parametrize("field", ["structured", "nanovdb", "tets"])
def test_raycast_gradient_shading(pkg, oracle, gpu_ctx, field):
    wls = pkg.workloads
    if field == "structured":
        wl = wls.blobs_raycast(n=48, width=96, height=80)
    elif field == "nanovdb":
        wl = dataclasses.replace(
            wls.cloud_nanovdb_pathtrace(n=64, width=80, height=48, frames=1, brick=8),
            renderer="raycast")
    else:
        wl = wls.tets_raycast(n=10)
    plain, _ = _raycast_pair(pkg, oracle, gpu_ctx, wl)
    g, o = _raycast_pair(pkg, oracle, gpu_ctx, _shaded(wl))
    err = np.abs(g - o).max()
    print(field, "gradient raycast max-abs", err)
    assert err <= 1e-3
    assert np.abs(g - plain).max() > 1e-2  # the shading does change the image


def test_raycast_gradient_quad_light(pkg, oracle, gpu_ctx):
    """A quad light sampled per sample from the pixel's stream (identical
    draws on both sides; the reference's ray caster leaves its generator
    unseeded, so these draws are not the reference's)."""
    wl = pkg.workloads.blobs_raycast(n=48, width=96, height=80)
    quad = dict(type="quad", position=(-0.3, 0.8, -0.3), edge1=(0.6, 0, 0), edge2=(0, 0, 0.6),
                intensity=3.0, color=(1, 1, 1), visible=True)
    g, o = _raycast_pair(pkg, oracle, gpu_ctx, _shaded(wl, lights=(SUN, quad)))
    err = np.abs(g - o).max()
    print("gradient raycast with quad light max-abs", err)
    assert err <= 1e-3


@pytest.mark.parametrize("ambient_samples", [0, 1])
def test_pathtrace_gradient_shading(pkg, oracle, gpu_ctx, ambient_samples):
    wl = pkg.workloads.fbm_pathtrace(n=32, width=64, height=48, frames=1, brick=8)
    wl = _shaded(wl, ambient_samples=ambient_samples)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, gbuffer=True)
    ow = oracle.OracleWorkload(wl)
    gs.frame.reset()
    gs.render(0)
    gpu_ctx.sync()
    agree = oracle.stream_agreement(gs.frame.color_array(), ow.pathtrace_frame(0)[0])
    print("gradient pathtrace stream agreement", agree)
    assert agree >= oracle.STREAM_AGREEMENT_MIN
    # the g-buffer normal is the gradient normal (or the AO draw where there is none)
    n_gpu = gs.frame.gbuffer_arrays()["normal"]
    n_orc = ow.gbuffer()["normal"]
    close = np.isclose(n_gpu, n_orc, rtol=0, atol=1e-5).all(-1)
    print("normal agreement", close.mean())
    assert close.mean() >= oracle.STREAM_AGREEMENT_MIN
    hit = np.linalg.norm(n_orc, axis=-1) > 0
    assert hit.mean() > 0.2


@pytest.mark.parametrize("gradient", [False, True])
def test_raycast_with_surfaces(pkg, oracle, gpu_ctx, gradient):
    """Ray caster over the mixed scene: matte surfaces under a directional and
    a quad light, the volume in front of them, the quad as backdrop."""
    wl = pkg.workloads.mixed_pathtrace(level=3, vol_n=24, width=64, height=48, frames=1)
    sun = dict(type="directional", direction=(0.4, -0.8, -0.3), color=(1, 1, 1), intensity=1.5)
    wl = dataclasses.replace(wl, renderer="raycast", lights=list(wl.lights) + [sun],
                             params=dict(volume_sampling_rate=1.0, gradient_shading=gradient))
    g, o = _raycast_pair(pkg, oracle, gpu_ctx, wl)
    err = np.abs(g - o).max()
    print("raycast with surfaces max-abs", err, "gradient", gradient)
    assert err <= 1e-3
    assert (g[..., 3] > 0.99).mean() > 0.5  # the sphere covers most of the image
