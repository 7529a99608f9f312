"""G-buffer channels on the GPU (depth, normal, albedo, primitive / object /
instance id; Frame::fillGBuffer, DeviceCopyableObjects.h:2745-2774) against
the oracle's, frame by frame over accumulated frames, for surface, mixed,
tetrahedral, instanced, NanoVDB and ray-caster scenes."""
import ctypes as C
import dataclasses
import math

import numpy as np
import pytest

import test_gbuffer_oracle as G

pytestmark = pytest.mark.gpu

KEYS = ("primitive_id", "object_id", "instance_id")


def _compare(gpu, orc, min_agree=0.995):
    """Ids equal and depth / normal / albedo within float rounding on the
    pixels whose random streams agree (all of them in practice)."""
    same = np.ones(gpu["depth"].shape, bool)
    for k in KEYS:
        same &= gpu[k] == orc[k]
    d_ok = np.isclose(gpu["depth"], orc["depth"], rtol=1e-5, atol=0)
    n_ok = np.isclose(gpu["normal"], orc["normal"], rtol=0, atol=1e-5).all(-1)
    a_ok = np.isclose(gpu["albedo"], orc["albedo"], rtol=0, atol=1e-6).all(-1)
    agree = same & d_ok & n_ok & a_ok
    print("g-buffer agreement", agree.mean(), "ids", same.mean(), "depth", d_ok.mean())
    assert agree.mean() >= min_agree


def _gpu_frames(gs, frames):
    gs.frame.reset()
    out = []
    for a in range(frames):
        gs.render(a)
        gs.ctx.sync()
        g = gs.frame.gbuffer_arrays()
        g["depth"] = gs.frame.depth_array()
        out.append(g)
    return out


def _set_ids(gs, ow, vol_ids=None, surf_ids=None):
    if vol_ids is not None:
        gs.scene.volumes[0].object_id, gs.scene.volumes[0].instance_id = vol_ids
        ow.scene.varr[0].object_id, ow.scene.varr[0].instance_id = vol_ids
    for k, ids in enumerate(surf_ids or []):
        gs.scene.surfaces[k].object_id, gs.scene.surfaces[k].instance_id = ids
        ow.scene.marr[k].object_id, ow.scene.marr[k].instance_id = ids


def _check(pkg, oracle, gpu_ctx, wl, frames=3, volume=None, vol_ids=None, surf_ids=None,
           xfm=None):
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, volume=volume, gbuffer=True)
    ow = oracle.OracleWorkload(wl)
    if xfm is not None:
        ow.scene.varr[0].xfm = (C.c_float * 12)(*np.asarray(xfm, np.float32).ravel())
    _set_ids(gs, ow, vol_ids, surf_ids)
    got = _gpu_frames(gs, frames)
    for a in range(frames):
        _compare(got[a], ow.gbuffer(accum_ids=tuple(range(a + 1))))
    return got[-1]


@pytest.mark.parametrize("spp", [1, 4])
def test_surface_quad(pkg, oracle, gpu_ctx, spp):
    wl = G._surface_workload(pkg, tilt=0.7, spp=spp)
    g = _check(pkg, oracle, gpu_ctx, wl, surf_ids=[(7, 3)])
    hit = g["depth"] < 1e30
    assert hit.mean() > 0.2 and (g["object_id"][hit] == 7).all()


def test_mixed_scene(pkg, oracle, gpu_ctx):
    wl = pkg.workloads.mixed_pathtrace(level=3, vol_n=24, width=48, height=32, frames=1)
    g = _check(pkg, oracle, gpu_ctx, wl, vol_ids=(21, 22), surf_ids=[(31, 32)])
    kinds = set(np.unique(g["object_id"]).tolist())
    print("object ids present", kinds)
    assert {21, 31} <= kinds  # both the volume and the sphere appear


def test_tets_skipped_cell_primitive_ids(pkg, oracle, gpu_ctx):
    wl, idx = G._tets_workload(pkg)
    g = _check(pkg, oracle, gpu_ctx, wl, vol_ids=(11, 12))
    hit = g["depth"] < 1e30
    assert hit.sum() > 200 and g["primitive_id"][hit].min() >= 1


def test_instanced_cube_normal_transform(pkg, oracle, gpu_ctx):
    wl = G._cube_workload(pkg)
    xfm = np.array([[0, -1, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0]], np.float32)
    vol = pkg.vrt.DeviceVolume(gpu_ctx, wl.voxels, wl.origin, wl.spacing, wl.tf,
                               wl.value_range, wl.unit_distance, wl.grid_dims, xfm=xfm)
    g = _check(pkg, oracle, gpu_ctx, wl, volume=vol, xfm=xfm)
    hit = g["depth"] < 1e30
    np.testing.assert_allclose(np.linalg.norm(g["normal"][hit], axis=1), 1.0, atol=1e-5)


def test_nanovdb_cloud(pkg, oracle, gpu_ctx):
    wl = pkg.workloads.cloud_nanovdb_pathtrace(n=64, width=64, height=36, frames=1, brick=8)
    g = _check(pkg, oracle, gpu_ctx, wl, vol_ids=(41, 42))
    hit = g["depth"] < 1e30
    assert hit.sum() > 100 and (g["primitive_id"][hit] == 0).all()


def test_raycast_two_volumes(pkg, oracle, gpu_ctx):
    vrt, wls = pkg.vrt, pkg.workloads
    W, H = 48, 24
    vox = np.full((4, 4, 4), 0.5, np.float32)
    tf = wls.ramp_tf(8)
    origins = ((-1.0, -0.45, -0.45), (0.1, -0.45, -0.45))
    dv = [vrt.DeviceVolume(gpu_ctx, vox, o, (0.3, 0.3, 0.3), tf, (0.0, 1.0), 0.5)
          for o in origins]
    ov = [oracle.OracleVolume(vox, o, (0.3, 0.3, 0.3), tf, (0.0, 1.0), 0.5) for o in origins]
    scene = vrt.make_scene(dv)
    osc = oracle.OracleScene(ov)
    for k, ids in enumerate(((5, 50), (6, 60))):
        scene.volumes[k].object_id, scene.volumes[k].instance_id = ids
        osc.varr[k].object_id, osc.varr[k].instance_id = ids
    pos = (0.0, 0.0, 3.0)
    cam_args = dict(position=pos, direction=wls.look_at(pos), up=(0, 1, 0),
                    fovy=math.radians(50.0), aspect=W / H)
    frame = vrt.DeviceFrame(gpu_ctx, W, H, gbuffer=True)
    vrt.raycast(gpu_ctx, scene, vrt.camera(**cam_args), vrt.params(volume_sampling_rate=1.0),
                frame, 0)
    gpu_ctx.sync()
    g = frame.gbuffer_arrays()
    g["depth"] = frame.depth_array()
    o = {"depth": np.full((H, W), 1e31, np.float32), "normal": np.zeros((H, W, 3), np.float32),
         "albedo": np.zeros((H, W, 3), np.float32)}
    for k in KEYS:
        o[k] = np.zeros((H, W), np.uint32)
    oracle.lib().orc_raycast_gbuffer(
        C.byref(osc.s), C.byref(oracle.camera(**cam_args)),
        C.byref(oracle.params(volume_sampling_rate=1.0)), W, H,
        *[o[k].ctypes.data for k in ("depth", "normal", "albedo") + KEYS], 1)
    for k in ("depth",) + KEYS:
        np.testing.assert_array_equal(g[k], o[k])
    assert not g["normal"].any() and not g["albedo"].any()
    assert {5, 6, G.NO_ID} <= set(np.unique(g["object_id"]).tolist())


def test_radiance_cache_frame_keeps_first_hits(pkg, gpu_ctx):
    """The cache replaces the path after its second vertex: the first hits,
    and so the g-buffer, are those of the plain path tracer."""
    wl = pkg.workloads.cloud_nanovdb_pathtrace(n=64, width=64, height=36, frames=1, brick=8)
    plain = pkg.workloads.GpuScene(gpu_ctx, wl, gbuffer=True)
    ref = _gpu_frames(plain, 2)[-1]
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, volume=plain.volume, gbuffer=True)
    nrc = pkg.vrt.Nrc(gpu_ctx)
    try:
        gs.frame.reset()
        opts = gs.nrc_options()
        for a in range(2):
            gs.render_nrc(nrc, opts, a)
        gpu_ctx.sync()
        g = gs.frame.gbuffer_arrays()
        g["depth"] = gs.frame.depth_array()
        for k in ("depth",) + KEYS:
            np.testing.assert_array_equal(g[k], ref[k])
        np.testing.assert_array_equal(g["normal"], ref["normal"])
    finally:
        nrc.close()


def test_gbuffer_needs_depth(pkg, gpu_ctx):
    vrt = pkg.vrt
    wl = G._surface_workload(pkg)
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, gbuffer=True)
    gs.frame.desc.depth = None
    with pytest.raises(vrt.VrtError, match="depth"):
        gs.render(0)


def test_raycast_surfaces_gbuffer(pkg, oracle, gpu_ctx):
    """Ray caster with surfaces: the first surface's depth, normal, albedo and
    ids, then the volume box's ids over them (Raycast_impl.cpp:135-143,
    230-237)."""
    wl = pkg.workloads.mixed_pathtrace(level=3, vol_n=24, width=48, height=32, frames=1)
    wl = dataclasses.replace(wl, renderer="raycast", params=dict(volume_sampling_rate=1.0))
    g = _check(pkg, oracle, gpu_ctx, wl, frames=2, vol_ids=(21, 22), surf_ids=[(31, 32)])
    assert (g["depth"] < 1e30).mean() > 0.5
