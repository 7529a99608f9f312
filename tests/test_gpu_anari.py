"""The ANARI front end (include/anari_vrt.h) on the GPU: scenes built through
the ANARI C API the way an application builds them render exactly the
images of the vrt C ABI for the same inputs (every field type, surfaces,
quad lights, both renderers); instance transforms agree with the oracle;
accumulation restarts on commits (Frame.cpp:499-513); errors reach the
status callback."""
import ctypes as C
import dataclasses
import math

import numpy as np
import pytest

import anari_api as A

pytestmark = pytest.mark.gpu


def _tf_table(pkg, colors, opacity):
    """The table the front end builds from these arrays (vrt_tf_build)."""
    L = pkg.vrt.lib()
    L.vrt_tf_build.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                               C.c_void_p]
    col = np.ascontiguousarray(colors, np.float32)
    op = np.ascontiguousarray(opacity, np.float32)
    uni = np.ones(4, np.float32)
    n = L.vrt_tf_build(col.ctypes.data, len(col), col.shape[1], op.ctypes.data, len(op),
                       uni.ctypes.data, None)
    out = np.empty((n, 4), np.float32)
    L.vrt_tf_build(col.ctypes.data, len(col), col.shape[1], op.ctypes.data, len(op),
                   uni.ctypes.data, out.ctypes.data)
    return out


def _anari_scene(a, wl, xfm=None, bvh_builder=None):
    """World, camera, renderer and frame of a workloads.Workload; returns
    (frame, objects by name)."""
    objs = {}
    vols = []
    if wl.voxels is not None or wl.nanovdb is not None or wl.tets is not None:
        if wl.voxels is not None:
            f = a.new(A.SPATIAL_FIELD, "structuredRegular")
            a.set(f, "data", A.ARRAY3D, a.array3d(wl.voxels))
            a.set(f, "origin", A.FLOAT32_VEC3, wl.origin)
            a.set(f, "spacing", A.FLOAT32_VEC3, wl.spacing)
        elif wl.nanovdb is not None:
            f = a.new(A.SPATIAL_FIELD, "nanovdb")
            a.set(f, "data", A.ARRAY1D, a.array1d(wl.nanovdb, A.UINT8))
            a.set(f, "filter", A.STRING, wl.filter)
        else:
            t = wl.tets
            f = a.new(A.SPATIAL_FIELD, "unstructured")
            ntet = len(t["index"])
            a.set(f, "vertex.position", A.ARRAY1D, a.array1d(t["positions"], A.FLOAT32_VEC3))
            a.set(f, "vertex.data", A.ARRAY1D, a.array1d(t["vertex_data"], A.FLOAT32))
            a.set(f, "index", A.ARRAY1D, a.array1d(np.asarray(t["index"], np.uint32).ravel(),
                                                   A.UINT32))
            a.set(f, "cell.index", A.ARRAY1D,
                  a.array1d(np.arange(ntet, dtype=np.uint32) * 4, A.UINT32))
            a.set(f, "cell.type", A.ARRAY1D, a.array1d(np.full(ntet, 10, np.uint8), A.UINT8))
        v = a.new(A.VOLUME, "transferFunction1D")
        a.set(v, "value", A.SPATIAL_FIELD, f)
        a.set(v, "color", A.ARRAY1D, a.array1d(wl.tf[:, :3], A.FLOAT32_VEC3))
        a.set(v, "opacity", A.ARRAY1D, a.array1d(wl.tf[:, 3], A.FLOAT32))
        a.set(v, "valueRange", A.FLOAT32_BOX1, wl.value_range)
        a.set(v, "unitDistance", A.FLOAT32, wl.unit_distance)
        if wl.grid_dims is not None:
            a.set(v, "majorantGrid", A.UINT32_VEC3, wl.grid_dims)
        a.commit(f, v)
        vols.append(v)
        objs.update(field=f, volume=v)
    lights = []
    for L in wl.lights:
        if L.get("type") == "hdri":
            l = a.new(A.LIGHT, "hdri")
            a.set(l, "radiance", A.ARRAY2D, a.array2d(L["radiance"], A.FLOAT32_VEC3))
            a.set(l, "up", A.FLOAT32_VEC3, L.get("up", (1, 0, 0)))
            a.set(l, "direction", A.FLOAT32_VEC3, L.get("direction", (0, 0, 1)))
            a.set(l, "scale", A.FLOAT32, L.get("scale", 1.0))
            a.set(l, "visible", A.BOOL, int(L.get("visible", True)))
            a.commit(l)
            lights.append(l)
            continue
        if L.get("type", "directional") == "quad":
            l = a.new(A.LIGHT, "quad")
            for k in ("position", "edge1", "edge2"):
                a.set(l, k, A.FLOAT32_VEC3, L[k])
            a.set(l, "intensity", A.FLOAT32, L["intensity"])
        else:
            l = a.new(A.LIGHT, "directional")
            a.set(l, "direction", A.FLOAT32_VEC3, L["direction"])
            a.set(l, "irradiance", A.FLOAT32, L["intensity"])
        a.set(l, "color", A.FLOAT32_VEC3, L["color"])
        a.set(l, "visible", A.BOOL, int(L.get("visible", True)))
        a.commit(l)
        lights.append(l)
    surfs = []
    for m in wl.meshes or []:
        g = a.new(A.GEOMETRY, "triangle")
        a.set(g, "vertex.position", A.ARRAY1D, a.array1d(m["vertices"], A.FLOAT32_VEC3))
        if bvh_builder is not None:
            a.set(g, "bvhBuilder", A.STRING, bvh_builder)
        objs.setdefault("geometries", []).append(g)
        if m.get("indices") is not None:
            a.set(g, "primitive.index", A.ARRAY1D, a.array1d(m["indices"], A.UINT32_VEC3))
        mat = a.new(A.MATERIAL, "matte")
        a.set(mat, "color", A.FLOAT32_VEC3, m["color"])
        s = a.new(A.SURFACE)
        a.set(s, "geometry", A.GEOMETRY, g)
        a.set(s, "material", A.MATERIAL, mat)
        a.commit(g, mat, s)
        surfs.append(s)
        objs.setdefault("surfaces", []).append(s)
    insts = []
    for m in wl.instances or []:
        # a surface in a group placed by an instance transform
        g = a.new(A.GEOMETRY, "triangle")
        a.set(g, "vertex.position", A.ARRAY1D, a.array1d(m["vertices"], A.FLOAT32_VEC3))
        if m.get("indices") is not None:
            a.set(g, "primitive.index", A.ARRAY1D, a.array1d(m["indices"], A.UINT32_VEC3))
        mat = a.new(A.MATERIAL, "matte")
        a.set(mat, "color", A.FLOAT32_VEC3, m["color"])
        sf = a.new(A.SURFACE)
        a.set(sf, "geometry", A.GEOMETRY, g)
        a.set(sf, "material", A.MATERIAL, mat)
        if "object_id" in m:
            a.set(sf, "id", A.UINT32, m["object_id"])
        grp = a.new(A.GROUP)
        a.set(grp, "surface", A.ARRAY1D, a.objects([sf], A.SURFACE))
        inst = a.new(A.INSTANCE, "transform")
        a.set(inst, "group", A.GROUP, grp)
        m4 = np.eye(4, dtype=np.float32)
        m4[:3, :] = np.asarray(m["transform"], np.float32).reshape(3, 4)
        a.set(inst, "transform", A.FLOAT32_MAT4, m4.T.ravel())  # column-major
        if "instance_id" in m:
            a.set(inst, "id", A.UINT32, m["instance_id"])
        a.commit(g, mat, sf, grp, inst)
        insts.append(inst)
    world = a.new(A.WORLD)
    if insts:
        a.set(world, "instance", A.ARRAY1D, a.objects(insts, A.INSTANCE))
    if xfm is None:
        if vols:
            a.set(world, "volume", A.ARRAY1D, a.objects(vols, A.VOLUME))
    else:
        grp = a.new(A.GROUP)
        a.set(grp, "volume", A.ARRAY1D, a.objects(vols, A.VOLUME))
        inst = a.new(A.INSTANCE, "transform")
        a.set(inst, "group", A.GROUP, grp)
        m4 = np.eye(4, dtype=np.float32)
        m4[:3, :] = np.asarray(xfm, np.float32).reshape(3, 4)
        a.set(inst, "transform", A.FLOAT32_MAT4, m4.T.ravel())  # column-major
        a.commit(grp, inst)
        a.set(world, "instance", A.ARRAY1D, a.objects([inst], A.INSTANCE))
    if lights:
        a.set(world, "light", A.ARRAY1D, a.objects(lights, A.LIGHT))
    if surfs:
        a.set(world, "surface", A.ARRAY1D, a.objects(surfs, A.SURFACE))
    cam = a.new(A.CAMERA, "perspective")
    c = wl.camera
    a.set(cam, "position", A.FLOAT32_VEC3, c["position"])
    a.set(cam, "direction", A.FLOAT32_VEC3, c["direction"])
    a.set(cam, "up", A.FLOAT32_VEC3, c["up"])
    a.set(cam, "fovy", A.FLOAT32, c["fovy"])
    a.set(cam, "aspect", A.FLOAT32, c["aspect"])
    if "aperture_radius" in c:
        a.set(cam, "apertureRadius", A.FLOAT32, c["aperture_radius"])
        a.set(cam, "focusDistance", A.FLOAT32, c["focus_distance"])
    ren = a.new(A.RENDERER, wl.renderer)
    p = wl.params
    a.set(ren, "background", A.FLOAT32_VEC4, p.get("background", (0, 0, 0, 1)))
    for k, n, kind in (("ambient_color", "ambientColor", A.FLOAT32_VEC3),
                       ("ambient_radiance", "ambientRadiance", A.FLOAT32),
                       ("occlusion_distance", "ambientOcclusionDistance", A.FLOAT32),
                       ("ambient_samples", "ambientSamples", A.INT32),
                       ("max_bounce", "maxBounce", A.INT32),
                       ("pixel_samples", "pixelSamples", A.INT32),
                       ("volume_sampling_rate", "volumeSamplingRate", A.FLOAT32),
                       ("phase_g", "phaseG", A.FLOAT32),
                       ("gradient_shading", "gradientShading", A.BOOL),
                       ("ratio_tracking", "ratioTracking", A.BOOL)):
        if k in p:
            a.set(ren, n, kind, p[k])
    fr = a.new(A.FRAME)
    a.set(fr, "size", A.UINT32_VEC2, (wl.width, wl.height))
    a.set(fr, "channel.color", A.DATA_TYPE, A.FLOAT32_VEC4)
    a.set(fr, "world", A.WORLD, world)
    a.set(fr, "camera", A.CAMERA, cam)
    a.set(fr, "renderer", A.RENDERER, ren)
    a.commit(world, cam, ren, fr)
    objs.update(world=world, camera=cam, renderer=ren, frame=fr)
    return fr, objs


def _same_tf(pkg, wl):
    return dataclasses.replace(wl, tf=_tf_table(pkg, wl.tf[:, :3], wl.tf[:, 3]))


def _vrt_frames(pkg, ctx, wl, frames):
    gs = pkg.workloads.GpuScene(ctx, wl)
    gs.frame.reset()
    out = []
    for k in range(frames):
        gs.render(k)
        ctx.sync()
        out.append(gs.frame.color_array().copy())
    return out


def _lens(wl):
    return dataclasses.replace(wl, camera={**wl.camera, "aperture_radius": 0.06,
                                           "focus_distance": 1.8})


CASES = {
    "structured pathtrace": lambda w: w.fbm_pathtrace(n=32, width=64, height=48, frames=1,
                                                      brick=8, phase_g=0.3),
    "structured raycast": lambda w: w.blobs_raycast(n=48, width=64, height=48),
    "hdri mixed surfaces": lambda w: dataclasses.replace(
        w.mixed_pathtrace(level=3, vol_n=24, width=48, height=32, frames=1),
        lights=[dict(type="hdri", up=(0, 0, 1), direction=(1, 0, 0), scale=1.3,
                     radiance=np.random.default_rng(2).uniform(0.1, 2.0, (8, 16, 3))
                     .astype(np.float32))]),
    "raycast mixed surfaces": lambda w: dataclasses.replace(
        w.mixed_pathtrace(level=3, vol_n=24, width=48, height=32, frames=1),
        renderer="raycast", params=dict(volume_sampling_rate=1.0)),
    "gradient-shaded raycast": lambda w: dataclasses.replace(
        w.blobs_raycast(n=48, width=64, height=48),
        lights=[dict(type="directional", direction=(0.5, -0.7, -0.4), color=(1, 1, 1),
                     intensity=2.0)],
        params=dict(volume_sampling_rate=2.0, gradient_shading=True)),
    "gradient-shaded pathtrace": lambda w: dataclasses.replace(
        w.fbm_pathtrace(n=32, width=64, height=48, frames=1, brick=8),
        params=dict(max_bounce=4, ambient_samples=1, gradient_shading=True)),
    "nanovdb pathtrace": lambda w: w.cloud_nanovdb_pathtrace(n=64, width=64, height=36,
                                                            frames=1, brick=8),
    "tets pathtrace": lambda w: w.tets_pathtrace(n=8, width=48, height=32, frames=1, grid=8),
    "ratio-tracked pathtrace": lambda w: dataclasses.replace(
        w.fbm_pathtrace(n=32, width=64, height=48, frames=1, brick=8),
        params=dict(max_bounce=4, ambient_samples=1, ratio_tracking=True)),
    "mixed surfaces": lambda w: w.mixed_pathtrace(level=3, vol_n=24, width=48, height=32,
                                                  frames=1),
    "instanced surfaces pathtrace": lambda w: w.instanced_pathtrace(width=64, height=48),
    "instanced surfaces raycast": lambda w: w.instanced_pathtrace(width=64, height=48,
                                                                  renderer="raycast"),
    "thin-lens pathtrace": lambda w: _lens(w.fbm_pathtrace(n=32, width=64, height=48, frames=1,
                                                           brick=8)),
    "thin-lens raycast": lambda w: _lens(w.blobs_raycast(n=48, width=64, height=48)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_anari_matches_vrt(pkg, gpu_ctx, case):
    wl = _same_tf(pkg, CASES[case](pkg.workloads))
    frames = 1 if wl.renderer == "raycast" else 4
    ref = _vrt_frames(pkg, gpu_ctx, wl, frames)
    a = A.Anari()
    try:
        fr, _ = _anari_scene(a, wl)
        for k in range(frames):
            a.render(fr)
            img, t = a.map(fr)
            assert t == A.FLOAT32_VEC4 and img.shape == (wl.height, wl.width, 4)
            np.testing.assert_array_equal(img, ref[k])
        assert a.prop_i(fr, "numSamples") == frames
        assert a.prop_f(fr, "duration") > 0
        errors = [m for m in a.messages if m[0] <= 1]
        assert not errors, errors
        assert np.abs(ref[-1]).max() > 0
    finally:
        a.close()


def test_instance_transform_matches_oracle(pkg, oracle, gpu_ctx):
    """A volume placed by a rotated and translated instance: raycast within
    1e-3 of the oracle with the same transform."""
    wl = _same_tf(pkg, pkg.workloads.blobs_raycast(n=48, width=64, height=48))
    c, s = math.cos(0.4), math.sin(0.4)
    xfm = [[c, 0, s, 0.15], [0, 1, 0, -0.1], [-s, 0, c, 0.05]]
    a = A.Anari()
    try:
        fr, _ = _anari_scene(a, wl, xfm=xfm)
        a.render(fr)
        img, _ = a.map(fr)
    finally:
        a.close()
    ow = oracle.OracleWorkload(wl)
    ov = oracle.OracleVolume(wl.voxels, wl.origin, wl.spacing, wl.tf, wl.value_range,
                             wl.unit_distance, xfm=xfm)
    ow.scene = oracle.OracleScene([ov], [], [])
    o = ow.raycast()
    err = np.abs(img - o).max()
    print("instance transform raycast max-abs", err)
    assert err <= 1e-3
    assert np.abs(img - _vrt_frames(pkg, gpu_ctx, wl, 1)[0]).max() > 0.05  # it moved


def test_world_limits(pkg, gpu_ctx):
    """Five surfaces: the fifth beyond the surface slots renders as an
    identity instance (frames identical to the C ABI's scene of four
    surfaces and that instance); a fifth volume is reported as an error,
    not dropped silently."""
    base = pkg.workloads.instanced_pathtrace(width=48, height=36, renderer="raycast")
    meshes = [dict(vertices=m["vertices"] @ np.asarray(m["transform"])[:, :3].T
                   + np.asarray(m["transform"])[:, 3], indices=m["indices"], color=m["color"])
              for m in base.instances[:5]]
    for m in meshes:
        m["vertices"] = np.ascontiguousarray(m["vertices"], np.float32)
    ident = np.concatenate([np.eye(3), np.zeros((3, 1))], 1).astype(np.float32)
    wl = _same_tf(pkg, dataclasses.replace(base, instances=None, meshes=meshes))
    ref_wl = dataclasses.replace(wl, meshes=meshes[:4],
                                 instances=[dict(meshes[4], transform=ident)])
    ref = _vrt_frames(pkg, gpu_ctx, ref_wl, 1)[0]
    a = A.Anari()
    try:
        fr, objs = _anari_scene(a, wl)
        a.render(fr)
        img, _ = a.map(fr)
        np.testing.assert_array_equal(img, ref)
        assert not [m for m in a.messages if m[0] <= 1]
        # five volumes: the fifth is an error
        v = objs["volume"]
        a.set(objs["world"], "volume", A.ARRAY1D, a.objects([v] * 5, A.VOLUME))
        a.commit(objs["world"])
        a.render(fr)
        a.map(fr)
        errs = [m for m in a.messages if m[0] <= 1]
        assert errs and "more than 4 volumes" in errs[-1][-1], errs
    finally:
        a.close()


def test_accumulation_restarts_on_commit(pkg, gpu_ctx):
    wl = _same_tf(pkg, CASES["structured pathtrace"](pkg.workloads))
    a = A.Anari()
    try:
        fr, objs = _anari_scene(a, wl)
        for _ in range(3):
            a.render(fr)
        assert a.prop_i(fr, "numSamples") == 3
        a.render(fr)
        assert a.prop_i(fr, "numSamples") == 4
        # a camera change restarts accumulation: the image is frame 0 again
        cam = objs["camera"]
        p = np.array(wl.camera["position"]) * 1.01
        a.set(cam, "position", A.FLOAT32_VEC3, p)
        a.commit(cam)
        a.render(fr)
        assert a.prop_i(fr, "numSamples") == 1
        moved = dataclasses.replace(wl, camera={**wl.camera, "position": tuple(p.astype(np.float32))})
        img, _ = a.map(fr)
        np.testing.assert_array_equal(img, _vrt_frames(pkg, gpu_ctx, moved, 1)[0])
        # staged but uncommitted parameters change nothing
        a.set(objs["renderer"], "maxBounce", A.INT32, 0)
        a.render(fr)
        assert a.prop_i(fr, "numSamples") == 2
    finally:
        a.close()


def test_color_channel_types_and_errors(pkg, gpu_ctx):
    wl = _same_tf(pkg, pkg.workloads.blobs_raycast(n=32, width=40, height=30))
    a = A.Anari()
    try:
        fr, objs = _anari_scene(a, wl)
        a.render(fr)
        f32, _ = a.map(fr)
        a.set(fr, "channel.color", A.DATA_TYPE, A.UFIXED8_VEC4)
        a.commit(fr)
        a.render(fr)
        u8, t = a.map(fr)
        assert t == A.UFIXED8_VEC4 and u8.dtype == np.uint8
        np.testing.assert_array_equal(u8, (255 * np.clip(f32, 0, 1)).astype(np.uint8))
        # a frame without a world reports an error and renders nothing
        bad = a.new(A.FRAME)
        a.set(bad, "size", A.UINT32_VEC2, (8, 8))
        a.commit(bad)
        n = len(a.messages)
        a.render(bad)
        assert any(sev == 1 and "world" in msg for sev, _, msg in a.messages[n:])
        # an unknown spatial field subtype is reported when a frame uses it
        f = a.new(A.SPATIAL_FIELD, "amr")
        v = objs["volume"]
        a.set(v, "value", A.SPATIAL_FIELD, f)
        a.commit(f, v)
        n = len(a.messages)
        a.render(fr)
        assert any("amr" in msg for _, _, msg in a.messages[n:])
    finally:
        a.close()


def test_geometry_gpu_bvh_builder(pkg, gpu_ctx):
    """Extension parameter bvhBuilder = "lbvh" builds the triangle BVH on the
    GPU: the same images as the host SAH build, build time and node count as
    geometry properties, and a switch back to "sah" on re-commit."""
    wl = _same_tf(pkg, CASES["mixed surfaces"](pkg.workloads))
    ref = _vrt_frames(pkg, gpu_ctx, wl, 2)
    a = A.Anari()
    try:
        fr, objs = _anari_scene(a, wl, bvh_builder="lbvh")
        g = objs["geometries"][0]
        for k in range(2):
            a.render(fr)
            np.testing.assert_array_equal(a.map(fr)[0], ref[k])
        t_gpu, n_gpu = a.prop_f(g, "bvhBuildTime"), a.prop_i(g, "bvhNodes")
        assert t_gpu is not None and t_gpu > 0 and n_gpu > 0
        a.set(g, "bvhBuilder", A.STRING, "sah")
        a.commit(g)
        a.render(fr)
        np.testing.assert_array_equal(a.map(fr)[0], ref[0])  # accumulation restarted
        assert a.prop_i(fr, "numSamples") == 1
        assert a.prop_i(g, "bvhNodes") > 0
        print("lbvh", t_gpu * 1e3, "ms", n_gpu, "nodes; sah", a.prop_f(g, "bvhBuildTime") * 1e3,
              "ms", a.prop_i(g, "bvhNodes"), "nodes")
        errors = [m for m in a.messages if m[0] <= 1]
        assert not errors, errors
    finally:
        a.close()


GBUF = (("channel.normal", A.FLOAT32_VEC3, "normal"), ("channel.albedo", A.FLOAT32_VEC3, "albedo"),
        ("channel.primitiveId", A.UINT32, "primitive_id"),
        ("channel.objectId", A.UINT32, "object_id"),
        ("channel.instanceId", A.UINT32, "instance_id"))


def test_gbuffer_channels_match_abi(pkg, gpu_ctx):
    """channel.depth/normal/albedo/primitiveId/objectId/instanceId with the
    "id" parameters of surface, volume and world (the world's own objects sit
    under an instance with the world's id, World.cpp:96): bit-identical to
    the C ABI frame's g-buffer."""
    wl = _same_tf(pkg, CASES["mixed surfaces"](pkg.workloads))
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, gbuffer=True)
    gs.scene.volumes[0].object_id, gs.scene.volumes[0].instance_id = 21, 99
    gs.scene.surfaces[0].object_id, gs.scene.surfaces[0].instance_id = 31, 99
    gs.frame.reset()
    for k in range(2):
        gs.render(k)
    gpu_ctx.sync()
    ref = gs.frame.gbuffer_arrays()
    ref["depth"] = gs.frame.depth_array()
    a = A.Anari()
    try:
        fr, objs = _anari_scene(a, wl)
        a.set(objs["volume"], "id", A.UINT32, 21)
        a.set(objs["surfaces"][0], "id", A.UINT32, 31)
        a.set(objs["world"], "id", A.UINT32, 99)
        a.set(fr, "channel.depth", A.DATA_TYPE, A.FLOAT32)
        for name, t, _ in GBUF:
            a.set(fr, name, A.DATA_TYPE, t)
        a.commit(objs["volume"], objs["surfaces"][0], objs["world"], fr)
        for k in range(2):
            a.render(fr)
        depth, t = a.map(fr, "channel.depth")
        assert t == A.FLOAT32
        np.testing.assert_array_equal(depth, ref["depth"])
        for name, t, key in GBUF:
            img, tt = a.map(fr, name)
            assert tt == t
            np.testing.assert_array_equal(img, ref[key])
        assert {21, 31} <= set(np.unique(ref["object_id"]).tolist())
        assert set(np.unique(ref["instance_id"]).tolist()) <= {99, 0xFFFFFFFF}
        errors = [m for m in a.messages if m[0] <= 1]
        assert not errors, errors
    finally:
        a.close()


def test_gbuffer_channels_need_depth(pkg, gpu_ctx):
    """Without channel.depth the reference never fills the other channels
    (Frame::fillGBuffer returns early): they map as zeros."""
    wl = _same_tf(pkg, CASES["mixed surfaces"](pkg.workloads))
    a = A.Anari()
    try:
        fr, _ = _anari_scene(a, wl)
        for name, t, _ in GBUF:
            a.set(fr, name, A.DATA_TYPE, t)
        a.commit(fr)
        a.render(fr)
        for name, t, _ in GBUF:
            img, tt = a.map(fr, name)
            assert tt == t and not img.any()
        assert a.map(fr, "channel.depth")[0] is None
    finally:
        a.close()


def test_frame_completion_callback(pkg, gpu_ctx):
    """frameCompletionCallback runs once per rendered frame, after the frame
    is complete, with the frame handle and the user data; a callback may
    release its own frame."""
    import threading
    import time
    wl = _same_tf(pkg, CASES["structured raycast"](pkg.workloads))
    a = A.Anari()
    calls = []
    done = threading.Event()

    def on_frame(user, dev, frame):
        calls.append((user, frame, threading.get_ident()))
        if len(calls) == 3:
            done.set()

    cb = A.FRAME_CALLBACK(on_frame)
    try:
        fr, objs = _anari_scene(a, wl)
        a.set(fr, "frameCompletionCallback", A.FRAME_COMPLETION_CALLBACK,
              C.cast(cb, C.c_void_p).value)
        a.set(fr, "frameCompletionCallbackUserData", A.VOID_POINTER, 1234)
        a.commit(fr)
        for _ in range(3):
            a.render(fr, wait=False)
        assert done.wait(30.0), calls
        time.sleep(0.1)
        assert len(calls) == 3
        assert all(u == 1234 and f == fr for u, f, _ in calls)
        assert all(tid != threading.get_ident() for _, _, tid in calls)
        img, _ = a.map(fr)
        assert np.abs(img).max() > 0

        # a second frame whose callback releases it
        released = threading.Event()
        fr2 = a.new(A.FRAME)

        def release_self(user, dev, frame):
            a.L.anariRelease(a.dev, frame)
            released.set()

        cb2 = A.FRAME_CALLBACK(release_self)
        a.set(fr2, "size", A.UINT32_VEC2, (wl.width, wl.height))
        a.set(fr2, "channel.color", A.DATA_TYPE, A.FLOAT32_VEC4)
        for k in ("world", "camera", "renderer"):
            a.set(fr2, k, {"world": A.WORLD, "camera": A.CAMERA, "renderer": A.RENDERER}[k],
                  objs[k])
        a.set(fr2, "frameCompletionCallback", A.FRAME_COMPLETION_CALLBACK,
              C.cast(cb2, C.c_void_p).value)
        a.commit(fr2)
        a.render(fr2, wait=False)
        assert released.wait(30.0)
        errors = [m for m in a.messages if m[0] <= 1]
        assert not errors, errors
    finally:
        a.close()


def test_device_released_inside_callback(pkg, gpu_ctx):
    """The last references -- the frame's, then the device's -- released from
    inside the frame's completion callback: the device is destroyed on the
    completion thread while that callback runs, and the thread then ends
    without touching the freed device; devices created afterwards work."""
    import threading
    import time
    a = A.Anari()
    made = []

    def keep(h):
        made.append(h)
        return h

    f = keep(a.new(A.SPATIAL_FIELD, "structuredRegular"))
    a.set(f, "data", A.ARRAY3D, keep(a.array3d(np.full((8, 8, 8), 0.5, np.float32))))
    a.set(f, "origin", A.FLOAT32_VEC3, (-0.5, -0.5, -0.5))
    a.set(f, "spacing", A.FLOAT32_VEC3, (1 / 7, 1 / 7, 1 / 7))
    v = keep(a.new(A.VOLUME, "transferFunction1D"))
    a.set(v, "value", A.SPATIAL_FIELD, f)
    a.set(v, "color", A.ARRAY1D, keep(a.array1d(np.ones((2, 3), np.float32), A.FLOAT32_VEC3)))
    a.set(v, "opacity", A.ARRAY1D, keep(a.array1d(np.array([0, 1], np.float32), A.FLOAT32)))
    a.commit(f, v)
    world = keep(a.new(A.WORLD))
    a.set(world, "volume", A.ARRAY1D, keep(a.objects([v], A.VOLUME)))
    cam = keep(a.new(A.CAMERA, "perspective"))
    a.set(cam, "position", A.FLOAT32_VEC3, (0, 0, 3))
    a.set(cam, "direction", A.FLOAT32_VEC3, (0, 0, -1))
    a.set(cam, "up", A.FLOAT32_VEC3, (0, 1, 0))
    ren = keep(a.new(A.RENDERER, "raycast"))
    a.commit(world, cam, ren)
    fr = a.new(A.FRAME)
    a.set(fr, "size", A.UINT32_VEC2, (32, 32))
    a.set(fr, "channel.color", A.DATA_TYPE, A.FLOAT32_VEC4)
    a.set(fr, "world", A.WORLD, world)
    a.set(fr, "camera", A.CAMERA, cam)
    a.set(fr, "renderer", A.RENDERER, ren)
    done = threading.Event()

    def last_callback(user, dev, frame):
        a.L.anariRelease(dev, frame)
        a.L.anariRelease(dev, dev)  # the device's last reference
        done.set()

    cb = A.FRAME_CALLBACK(last_callback)
    a.set(fr, "frameCompletionCallback", A.FRAME_COMPLETION_CALLBACK,
          C.cast(cb, C.c_void_p).value)
    a.commit(fr)
    a.release(*made)  # everything but the frame and the device
    a.render(fr, wait=False)
    assert done.wait(30.0)
    a.dev = None  # released by the callback
    time.sleep(0.2)  # the completion thread leaves its loop
    errors = [m for m in a.messages if m[0] <= 1]
    assert not errors, errors
    b = A.Anari()
    try:
        fr2, _ = _anari_scene(b, _same_tf(pkg, CASES["structured raycast"](pkg.workloads)))
        b.render(fr2)
        img, _ = b.map(fr2)
        assert np.abs(img).max() > 0
    finally:
        b.close()


def test_taa_frames_match_abi(pkg, gpu_ctx):
    """Renderer "taa" / "taaAlpha" with the camera moved between frames by
    commits: the resolved frames are those of the C ABI with the previous
    camera passed along."""
    vrt = pkg.vrt
    wl = pkg.workloads.fbm_pathtrace(n=24, width=40, height=32, frames=1, brick=8)
    wl = _same_tf(pkg, dataclasses.replace(
        wl, params=dict(max_bounce=3, ambient_samples=1, taa=True, taa_alpha=0.25)))
    poses = [np.array(wl.camera["position"]) + 0.015 * k * np.array([1.0, 0.4, 0.0])
             for k in range(3)]
    gs = pkg.workloads.GpuScene(gpu_ctx, wl, taa=True)
    ref, prev = [], None
    for k, p in enumerate(poses):
        c = dict(wl.camera, position=tuple(p))
        gs.camera = vrt.camera(**c)
        gs.params = vrt.params(**wl.params, prev_camera=vrt.camera(**prev) if prev else None)
        # every camera commit restarts the accumulation (accumID 0); the
        # TAA history carries over
        gs.frame.reset()
        gs.render(0)
        gpu_ctx.sync()
        ref.append(gs.frame.color_array().copy())
        prev = c
    a = A.Anari()
    try:
        fr, objs = _anari_scene(a, wl)
        a.set(objs["renderer"], "taa", A.BOOL, 1)
        a.set(objs["renderer"], "taaAlpha", A.FLOAT32, 0.25)
        a.commit(objs["renderer"])
        for k, p in enumerate(poses):
            a.set(objs["camera"], "position", A.FLOAT32_VEC3, p)
            a.commit(objs["camera"])
            a.render(fr)
            np.testing.assert_array_equal(a.map(fr)[0], ref[k])
        errors = [m for m in a.messages if m[0] <= 1]
        assert not errors, errors
    finally:
        a.close()
