"""G-buffer channels of the oracle (Frame::fillGBuffer,
DeviceCopyableObjects.h:2745-2774; Pathtrace_impl.cpp:226-285, 516-637;
Raycast_impl.cpp:18-20, 230-237): depth, geometric normal, albedo and
primitive / object / instance ids of the first radiance hit of a pixel's
closest sample, written where the depth passes the depth test.  Each check
is against the camera ray recomputed from the pixel's random stream, a
closed form, or an independent computation."""
import ctypes as C
import dataclasses
import math

import numpy as np
import pytest

NO_ID = 0xFFFFFFFF


def _quad_mesh(tilt=0.0, half=0.6):
    """Two triangles over [-half, half]^2 in the plane z = 0 rotated by `tilt`
    radians about x; triangle 0 below the diagonal y = x, 1 above it."""
    v = np.array([[-half, -half, 0], [half, -half, 0], [half, half, 0], [-half, half, 0]],
                 np.float64)
    c, s = math.cos(tilt), math.sin(tilt)
    v = v @ np.array([[1, 0, 0], [0, c, s], [0, -s, c]])
    return v.astype(np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _surface_workload(pkg, tilt=0.0, spp=1, width=48, height=40):
    wls = pkg.workloads
    v, i = _quad_mesh(tilt)
    pos = (0.1, 0.2, 2.0)
    return wls.Workload(
        name="gbuffer quad", voxels=None, origin=(0, 0, 0), spacing=(1, 1, 1),
        tf=wls.ramp_tf(4), value_range=(0.0, 1.0), unit_distance=1.0, grid_dims=None,
        camera=dict(position=pos, direction=wls.look_at(pos), up=(0, 1, 0),
                    fovy=math.radians(40.0), aspect=width / height),
        lights=[dict(type="directional", direction=(0.3, -0.5, -1.0), color=(1, 1, 1),
                     intensity=2.0)],
        params=dict(max_bounce=2, ambient_samples=0, pixel_samples=spp),
        width=width, height=height, renderer="pathtrace",
        meshes=[dict(vertices=v, indices=i, color=(0.25, 0.5, 0.75))])


def _oracle(oracle, wl, obj=7, inst=3):
    ow = oracle.OracleWorkload(wl)
    for k in range(len(ow.meshes)):
        ow.scene.marr[k].object_id = obj + k
        ow.scene.marr[k].instance_id = inst
    return ow


def _primary(oracle, ow, x, y, accum_id=0):
    """The first sample's camera ray: jitter = the first two numbers of the
    pixel's stream (Pathtrace_impl.cpp:510-520)."""
    W, H = ow.wl.width, ow.wl.height
    j = oracle.lcg_stream(x + W * y, accum_id, 2)
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    oracle.lib().orc_camera_ray(C.byref(ow.cam), x + j[0] - 0.5, y + j[1] - 0.5, W, H,
                                o.ctypes.data, d.ctypes.data)
    return o.astype(np.float64), d.astype(np.float64)


def test_surface_first_hit_closed_form(pkg, oracle):
    wl = _surface_workload(pkg)
    ow = _oracle(oracle, wl)
    g = ow.gbuffer()
    hits = misses = 0
    for y in range(wl.height):
        for x in range(wl.width):
            o, d = _primary(oracle, ow, x, y)
            t = -o[2] / d[2]
            P = o + t * d
            inside = max(abs(P[0]), abs(P[1])) < 0.6 - 1e-4
            outside = max(abs(P[0]), abs(P[1])) > 0.6 + 1e-4
            if inside and abs(P[0] - P[1]) > 1e-4:
                hits += 1
                assert g["depth"][y, x] == pytest.approx(t, rel=1e-5)
                np.testing.assert_array_equal(g["normal"][y, x], [0, 0, 1])  # faces the camera
                np.testing.assert_array_equal(g["albedo"][y, x],
                                              np.float32([0.25, 0.5, 0.75]))
                assert g["primitive_id"][y, x] == (0 if P[1] < P[0] else 1)
                assert (g["object_id"][y, x], g["instance_id"][y, x]) == (7, 3)
            elif outside:
                misses += 1
                assert g["depth"][y, x] == np.float32(1e31)
                assert g["primitive_id"][y, x] == g["object_id"][y, x] == NO_ID
                assert g["instance_id"][y, x] == NO_ID
                assert not g["normal"][y, x].any() and not g["albedo"][y, x].any()
    assert hits > 400 and misses > 300


def test_closest_sample_wins(pkg, oracle):
    """With 4 samples the g-buffer holds the closest one: sample 0 is the
    1-spp frame's sample (same stream), so its depth bounds the 4-spp depth
    from above, strictly wherever another sample lands nearer on the tilted
    plane."""
    g1 = _oracle(oracle, _surface_workload(pkg, tilt=0.7, spp=1)).gbuffer()
    g4 = _oracle(oracle, _surface_workload(pkg, tilt=0.7, spp=4)).gbuffer()
    hit = (g1["depth"] < 1e30) & (g4["depth"] < 1e30)
    assert hit.sum() > 400
    assert (g4["depth"][hit] <= g1["depth"][hit]).all()
    assert (g4["depth"][hit] < g1["depth"][hit]).mean() > 0.5
    # every channel comes from that one sample: a hit has the quad's values
    assert (g4["object_id"][hit] == 7).all() and (g4["primitive_id"][hit] <= 1).all()


def test_depth_test_across_frames(pkg, oracle):
    """Accumulated frames keep the nearest depth so far; a frame writes its
    channels where its depth is <= the stored one (ties: the later frame)."""
    wl = _surface_workload(pkg, tilt=0.7)
    ow = _oracle(oracle, wl)
    single = [ow.gbuffer(accum_ids=(a,)) for a in range(3)]
    acc = ow.gbuffer(accum_ids=(0, 1, 2))
    D = np.stack([s["depth"] for s in single])
    np.testing.assert_array_equal(acc["depth"], D.min(0))
    # the frame that wrote last: the last one whose depth equals the minimum
    last = 2 - np.argmax((D == D.min(0))[::-1], axis=0)
    for k in ("primitive_id", "object_id"):
        want = np.choose(last, [s[k] for s in single])
        np.testing.assert_array_equal(acc[k], want)


def _tets_workload(pkg, n=6):
    wl = pkg.workloads.tets_pathtrace(n=n, width=40, height=32, frames=1, grid=8)
    t = dict(wl.tets)
    idx = np.asarray(t["index"], np.uint32).reshape(-1, 4)
    T = len(idx)
    # cell 0: three indices, skipped by the reference (UnstructuredField.cpp:135-139);
    # tetrahedron k becomes cell k + 1
    t["index"] = np.concatenate([np.array([0, 1, 2], np.uint32), idx.ravel()])
    t["cell_index"] = np.concatenate([[0], 3 + 4 * np.arange(T)]).astype(np.uint32)
    t["cell_type"] = np.full(T + 1, 10, np.uint8)
    return dataclasses.replace(wl, tets=t, params=dict(wl.params, ambient_samples=0)), idx


def test_tet_primitive_ids_are_cell_indices(pkg, oracle):
    """A volume hit in an unstructured field reports the containing cell by
    its index in the ANARI cell arrays (elemID, UnstructuredField.cpp:146-151),
    skipped cells counted; the normal is 0 without ambient samples and the
    albedo the TF colour (0.8)."""
    wl, idx = _tets_workload(pkg)
    pos = np.asarray(wl.tets["positions"], np.float64)
    ow = oracle.OracleWorkload(wl)
    ow.scene.varr[0].object_id, ow.scene.varr[0].instance_id = 11, 12
    g = ow.gbuffer()
    hit = g["depth"] < 1e30
    assert hit.sum() > 200
    assert (g["object_id"][hit] == 11).all() and (g["instance_id"][hit] == 12).all()
    assert not g["normal"][hit].any()
    np.testing.assert_array_equal(g["albedo"][hit], np.float32(0.8))
    prim = g["primitive_id"][hit]
    assert prim.min() >= 1 and prim.max() <= len(idx)
    # the hit point lies in the reported cell (float64 barycentric coordinates)
    ok = 0
    ys, xs = np.nonzero(hit)
    for y, x in zip(ys, xs):
        o, d = _primary(oracle, ow, x, y)
        P = o + float(g["depth"][y, x]) * d
        v = pos[idx[g["primitive_id"][y, x] - 1]]
        b = np.linalg.solve((v[1:] - v[0]).T, P - v[0])
        ok += (b >= -1e-4).all() and b.sum() <= 1 + 1e-4
    assert ok >= 0.99 * len(ys)


def _cube_workload(pkg, width=40, height=32):
    wls = pkg.workloads
    n = 8
    pos = (0.3, 0.8, 2.2)
    tf = np.ones((4, 4), np.float32)  # alpha 1: the first tentative collision is accepted
    tf[:, :3] = (0.2, 0.4, 0.6)
    return wls.Workload(
        name="gbuffer cube", voxels=np.ones((n, n, n), np.float32), origin=(-0.5,) * 3,
        spacing=(1.0 / (n - 1),) * 3, tf=tf, value_range=(0.0, 1.0), unit_distance=0.3,
        grid_dims=None,
        camera=dict(position=pos, direction=wls.look_at(pos), up=(0, 1, 0),
                    fovy=math.radians(40.0), aspect=width / height),
        lights=[dict(type="directional", direction=(0.3, -0.5, -1.0), color=(1, 1, 1),
                     intensity=2.0)],
        params=dict(max_bounce=1, ambient_samples=1, pixel_samples=1),
        width=width, height=height, renderer="pathtrace")


def test_volume_normal_transform(pkg, oracle):
    """A constant cube about the origin rotated by 90 degrees about z is the
    same world-space medium, so every draw repeats; the random volume normal
    is then mapped by transpose(affineInv) = R (Instance.cpp:96-98,
    Pathtrace_impl.cpp:217-219) -- not by affineInv = R^T, not left as is."""
    wl = _cube_workload(pkg)
    ref = oracle.OracleWorkload(wl).gbuffer()
    ow = oracle.OracleWorkload(wl)
    ow.scene.varr[0].xfm = (C.c_float * 12)(0, -1, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0)
    rot = ow.gbuffer()
    hit = ref["depth"] < 1e30
    assert hit.sum() > 300
    np.testing.assert_array_equal(rot["depth"], ref["depth"])
    R = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    n_ref, n_rot = ref["normal"][hit], rot["normal"][hit]
    np.testing.assert_allclose(np.linalg.norm(n_ref, axis=1), 1.0, atol=1e-5)
    Rn = n_ref @ R.T
    same = np.all(n_rot == Rn, axis=1) | np.all(n_rot == -Rn, axis=1)
    assert same.all()
    # and faceforward: the recorded normal faces the viewer
    ys, xs = np.nonzero(hit)
    for y, x in list(zip(ys, xs))[:100]:
        _, d = _primary(oracle, ow, x, y)
        assert float(rot["normal"][y, x] @ -d) >= 0.0
    assert (rot["albedo"][hit] == np.float32([0.2, 0.4, 0.6])).all()
    assert (rot["primitive_id"][hit] == 0).all()  # structured field (:478-481)


def test_raycast_volume_box_ids(pkg, oracle):
    """Ray caster: the closest volume box's object and instance ids; depth
    1e31, primitive id unset, normal and albedo 0 (Raycast_impl.cpp:230-237)."""
    wls = pkg.workloads
    W, H = 48, 24
    pos = (0.0, 0.0, 3.0)
    vox = np.full((4, 4, 4), 0.5, np.float32)
    tf = wls.ramp_tf(8)
    vols = [oracle.OracleVolume(vox, org, (0.3, 0.3, 0.3), tf, (0.0, 1.0), 0.5)
            for org in ((-1.0, -0.45, -0.45), (0.1, -0.45, -0.45))]
    sc = oracle.OracleScene(vols)
    for k, (o_id, i_id) in enumerate(((5, 50), (6, 60))):
        sc.varr[k].object_id, sc.varr[k].instance_id = o_id, i_id
    cam = oracle.camera(pos, wls.look_at(pos), (0, 1, 0), math.radians(50.0), W / H)
    prm = oracle.params(volume_sampling_rate=1.0)
    g = {"depth": np.full((H, W), 1e31, np.float32), "normal": np.ones((H, W, 3), np.float32),
         "albedo": np.ones((H, W, 3), np.float32), "primitive_id": np.zeros((H, W), np.uint32),
         "object_id": np.zeros((H, W), np.uint32), "instance_id": np.zeros((H, W), np.uint32)}
    oracle.lib().orc_raycast_gbuffer(
        C.byref(sc.s), C.byref(cam), C.byref(prm), W, H,
        *[g[k].ctypes.data for k in ("depth", "normal", "albedo", "primitive_id", "object_id",
                                     "instance_id")], 1)
    assert (g["depth"] == np.float32(1e31)).all()
    assert (g["primitive_id"] == NO_ID).all()
    assert not g["normal"].any() and not g["albedo"].any()
    counts = {5: 0, 6: 0, NO_ID: 0}
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    for y in range(H):
        for x in range(W):
            oracle.lib().orc_camera_ray(C.byref(cam), x, y, W, H, o.ctypes.data, d.ctypes.data)
            want, margin = NO_ID, np.inf
            for k, lo in enumerate(((-1.0, -0.45, -0.45), (0.1, -0.45, -0.45))):
                lo = np.array(lo)
                hi = lo + 0.9
                with np.errstate(divide="ignore"):
                    t0, t1 = (lo - o) / d, (hi - o) / d
                tn, tf_ = np.minimum(t0, t1).max(), np.maximum(t0, t1).min()
                margin = min(margin, abs(tf_ - tn))
                if tf_ >= max(tn, 0):
                    want = (5, 6)[k]
            if margin < 1e-3:
                continue
            assert g["object_id"][y, x] == want
            assert g["instance_id"][y, x] == {5: 50, 6: 60, NO_ID: NO_ID}[want]
            counts[want] += 1
    assert min(counts.values()) > 50


def test_raycast_surface_matte_closed_form(pkg, oracle):
    """Ray caster over a quad facing the camera under one directional light:
    matte color/pi * E * max(0, n.L) + color * ambient, with n the geometric
    normal as stored (not faceforwarded, Raycast_impl.cpp:45-100); the
    g-buffer holds the surface's depth, normal, albedo and ids."""
    wl = dataclasses.replace(_surface_workload(pkg), renderer="raycast",
                             params=dict(ambient_radiance=0.25, ambient_color=(1, 1, 1)))
    L = np.array((0.3, -0.5, -1.0))
    ow = _oracle(oracle, wl)
    img = ow.raycast()
    g = ow.gbuffer()
    Ld = -L / np.linalg.norm(L)
    col = np.array([0.25, 0.5, 0.75])
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    hits = 0
    for y in range(wl.height):
        for x in range(wl.width):
            oracle.lib().orc_camera_ray(C.byref(ow.cam), x, y, wl.width, wl.height,
                                        o.ctypes.data, d.ctypes.data)
            t = -o[2] / d[2]
            P = o + t * d
            if max(abs(P[0]), abs(P[1])) > 0.6 - 1e-3:
                continue
            hits += 1
            # e1 x e2 of both triangles points to +z: towards the camera
            want = col / math.pi * 2.0 * max(0.0, Ld[2]) + col * 0.25
            np.testing.assert_allclose(img[y, x, :3], want, rtol=2e-6)
            assert img[y, x, 3] == 1.0
            assert g["depth"][y, x] == pytest.approx(t, rel=1e-5)
            np.testing.assert_allclose(g["normal"][y, x], [0, 0, 1], atol=1e-7)
            assert (g["object_id"][y, x], g["instance_id"][y, x]) == (7, 3)
    assert hits > 400
