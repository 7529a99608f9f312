"""Synthetic workloads of the benchmark configurations, as plain numpy data.

Each builder returns a ``Workload``: the field voxels (x-fastest, shape
(nz, ny, nx)), the transferFunction1D table and parameters, camera, lights,
renderer parameters and image size.  Both the GPU binding and the CPU
oracle consume the same object, so parity runs see bit-identical inputs.

blobs_raycast  -- raycast, 128^3 Gaussian blobs, 256-entry ramp TF, 512^2
fbm_pathtrace  -- pathtrace, 512^3 fBm, density scale 50, 32^3 majorant grid
                  (16^3-voxel macrocells), directional light + ambient, 1080p
cloud_nanovdb_pathtrace -- pathtrace, 1024^3 sparse fBm cloud as a NanoVDB grid
mixed_pathtrace -- pathtrace, 1.3 M-triangle icosphere + blob volume + quad light
tets_pathtrace -- pathtrace, 2 M tetrahedra (jittered 70^3 lattice), 4K
"""
import dataclasses
import math

import numpy as np

try:
    from . import vrt as _vrt
except ImportError:  # loaded by file path
    import importlib.util as _ilu
    import os as _os
    _spec = _ilu.spec_from_file_location(
        "vrt", _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "vrt.py"))
    _vrt = _ilu.module_from_spec(_spec)
    _spec.loader.exec_module(_vrt)


@dataclasses.dataclass
class Workload:
    name: str
    voxels: np.ndarray           # (nz, ny, nx) float32
    origin: tuple
    spacing: tuple
    tf: np.ndarray               # (n, 4) float32 table
    value_range: tuple
    unit_distance: float
    grid_dims: tuple             # majorant grid, None for none
    camera: dict                 # position, direction, up, fovy, aspect
    lights: list                 # dicts: type, direction, color, intensity
    params: dict                 # renderer parameters
    width: int
    height: int
    renderer: str                # "raycast" | "pathtrace"
    frames: int = 1
    nanovdb: np.ndarray = None   # NanoVDB grid buffer: a nanovdb field instead of voxels
    filter: str = "linear"       # nanovdb sampler
    meshes: list = None          # dicts: vertices (n,3), indices (m,3) or None, color
    tets: dict = None            # unstructured field: positions (V,3), index (T,4),
                                 # vertex_data (V,) [, cell_index, cell_type, cell_data]
    instances: list = None       # instanced meshes: dicts as `meshes` plus transform
                                 # (row-major 3x4 [L | t]) [, object_id, instance_id]

    @property
    def dims(self):
        nz, ny, nx = self.voxels.shape
        return (nx, ny, nz)


def ramp_tf(n=256):
    """Piecewise-linear RGBA ramp: rgb blue -> red, alpha 0 -> 1."""
    t = np.linspace(0.0, 1.0, n, dtype=np.float64)
    tf = np.stack([t, np.zeros_like(t), 1.0 - t, t], axis=1)
    return tf.astype(np.float32)


def albedo_tf(albedo=0.8, n=256):
    """Constant albedo, alpha = normalised value (density ramp)."""
    t = np.linspace(0.0, 1.0, n, dtype=np.float64)
    tf = np.stack([np.full_like(t, albedo)] * 3 + [t], axis=1)
    return tf.astype(np.float32)


def look_at(position, target=(0.0, 0.0, 0.0)):
    d = np.asarray(target, np.float64) - np.asarray(position, np.float64)
    return tuple(float(x) for x in d / np.linalg.norm(d))


def blobs_raycast(n=128, width=512, height=512, seed=0, threads=0):
    """Raycast: structuredRegular n^3 f32 Gaussian blobs (32 blobs), 256-entry
    ramp TF, perspective fovy 40 deg from distance 2.5, 1 spp, step 0.5 voxel."""
    vox = _vrt.gen_blobs((n, n, n), 32, seed, threads)
    sp = 1.0 / (n - 1)
    pos = (0.0, 0.0, 2.5)
    return Workload(
        name=f"raycast blobs {n}^3 {width}x{height}", voxels=vox, origin=(-0.5,) * 3,
        spacing=(sp,) * 3, tf=ramp_tf(256), value_range=(0.0, 1.0), unit_distance=sp,
        grid_dims=None,
        camera=dict(position=pos, direction=look_at(pos), up=(0, 1, 0),
                    fovy=math.radians(40.0), aspect=width / height),
        lights=[],
        params=dict(background=(0, 0, 0, 1), volume_sampling_rate=2.0, pixel_samples=1),
        width=width, height=height, renderer="raycast")


def fbm_pathtrace(n=512, width=1920, height=1080, frames=64, seed=1, brick=16,
                  density_scale=50.0, albedo=0.8, max_bounce=8, phase_g=0.0, threads=0):
    """Pathtrace: structuredRegular n^3 f32 fBm (6 octaves, lacunarity 2,
    gain 0.5) remapped to [0,1]; TF rgb = albedo, alpha = value with
    unitDistance = 1/density_scale; majorant grid of brick^3-voxel cells;
    one directional light plus the constant ambient term; Henyey-Greenstein
    phase of anisotropy phase_g (0: the reference's isotropic phase)."""
    vox = _vrt.gen_fbm((n, n, n), 6, 4.0, 2.0, 0.5, seed, threads)
    sp = 1.0 / (n - 1)
    g = max(1, (n + brick - 1) // brick)
    pos = (1.5, 1.0, 2.0)
    return Workload(
        name=f"pathtrace fbm {n}^3 {width}x{height}", voxels=vox, origin=(-0.5,) * 3,
        spacing=(sp,) * 3, tf=albedo_tf(albedo, 256), value_range=(0.0, 1.0),
        unit_distance=1.0 / density_scale, grid_dims=(g, g, g),
        camera=dict(position=pos, direction=look_at(pos), up=(0, 1, 0),
                    fovy=math.radians(40.0), aspect=width / height),
        lights=[dict(type="directional", direction=(-0.4, -1.0, -0.3), color=(1, 1, 1),
                     intensity=1.0)],
        params=dict(background=(0, 0, 0, 1), ambient_color=(1, 1, 1), ambient_radiance=0.5,
                    occlusion_distance=1e20, ambient_samples=1, max_bounce=max_bounce,
                    pixel_samples=1, phase_g=phase_g),
        width=width, height=height, renderer="pathtrace", frames=frames)


def cloud_nanovdb_pathtrace(n=1024, width=1920, height=1080, frames=256, seed=3,
                            active_fraction=0.05, albedo=0.8, max_bounce=8, brick=32,
                            unit_distance=None, phase_g=0.0, threads=0):
    """Sparse fBm cloud as a 'nanovdb' field (1024^3 index space, identity
    map, 8^3 leaves, ~5% active), path traced with a directional light and
    ambient term; majorant cells of `brick` voxels over the grid bounds."""
    buf, _ = _vrt.gen_nanovdb_cloud(n=n, seed=seed, active_fraction=active_fraction,
                                    threads=threads)
    ud = unit_distance if unit_distance is not None else 2.0 * n / 1024.0
    c = n / 2.0
    d = np.array([1.5, 1.0, 2.0]) / math.sqrt(7.25)
    pos = tuple(float(x) for x in c + 1.25 * n * d)  # 1.25 n from the centre
    bb = _grid_bounds(buf)
    g = tuple(max(1, int(math.ceil((bb[3 + a] - bb[a]) / brick))) for a in range(3))
    return Workload(
        name=f"pathtrace nanovdb cloud {n}^3 {width}x{height}", voxels=None, origin=None,
        spacing=None, tf=albedo_tf(albedo, 256), value_range=(0.0, 1.0), unit_distance=ud,
        grid_dims=g,
        camera=dict(position=pos, direction=look_at(pos, (c, c, c)), up=(0, 1, 0),
                    fovy=math.radians(40.0), aspect=width / height),
        lights=[dict(type="directional", direction=(-0.4, -1.0, -0.3), color=(1, 1, 1),
                     intensity=1.0)],
        params=dict(background=(0, 0, 0, 1), ambient_color=(1, 1, 1), ambient_radiance=0.5,
                    occlusion_distance=1e20, ambient_samples=1, max_bounce=max_bounce,
                    pixel_samples=1, phase_g=phase_g),
        width=width, height=height, renderer="pathtrace", frames=frames, nanovdb=buf)


def mixed_pathtrace(level=8, vol_n=256, width=1920, height=1080, frames=128, max_bounce=6,
                    seed=1, threads=0):
    """Mixed scene: a noise-displaced icosphere of 20 * 4^level triangles
    (level 8: 1,310,720), matte 0.7, enclosing a vol_n^3 Gaussian-blob volume
    (32 blobs, seed 1, ramp TF) and one quad light; camera and light inside
    the sphere; <= 6 bounces, 1 spp per frame, `frames` frames."""
    v, i = _vrt.gen_icosphere(level=level, radius=0.5, amp=0.08, freq=3.0, seed=2)
    vox = _vrt.gen_blobs((vol_n, vol_n, vol_n), count=32, seed=seed, threads=threads)
    sp = 0.5 / (vol_n - 1)
    pos = (0.08, 0.12, 0.36)
    g = max(1, (vol_n + 15) // 16)
    return Workload(
        name=f"pathtrace mixed icosphere {len(i)} tris + blobs {vol_n}^3 {width}x{height}",
        voxels=vox, origin=(-0.25,) * 3, spacing=(sp,) * 3, tf=ramp_tf(256),
        value_range=(0.0, 1.0), unit_distance=0.05, grid_dims=(g, g, g),
        camera=dict(position=pos, direction=look_at(pos), up=(0, 1, 0),
                    fovy=math.radians(60.0), aspect=width / height),
        lights=[dict(type="quad", position=(-0.08, 0.36, -0.08), edge1=(0.16, 0, 0),
                     edge2=(0, 0, 0.16), color=(1, 1, 1), intensity=12.0, visible=True)],
        params=dict(background=(0, 0, 0, 1), ambient_color=(1, 1, 1), ambient_radiance=0.0,
                    occlusion_distance=1e20, ambient_samples=0, max_bounce=max_bounce,
                    pixel_samples=1),
        width=width, height=height, renderer="pathtrace", frames=frames,
        meshes=[dict(vertices=v, indices=i, color=(0.7, 0.7, 0.7))])


def instance_transform(axis, angle, scale, translate):
    """Row-major 3x4 [L | t] with L = R(axis, angle) diag(scale)."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * (K @ K)
    L = R @ np.diag(np.asarray(scale, np.float64))
    return np.concatenate([L, np.asarray(translate, np.float64).reshape(3, 1)], 1).astype(np.float32)


def instanced_meshes(count=6, level=3, seed=4, spread=0.55):
    """`count` instances of one noise-displaced icosphere (20 * 4^level
    triangles): random axes and angles, non-uniform scales 0.08-0.2, placed
    on a ring of radius `spread`; instance ids 100 + i, object id 7."""
    v, i = _vrt.gen_icosphere(level=level, radius=1.0, amp=0.1, freq=3.0, seed=seed)
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        ang = 2 * math.pi * k / count
        t = (spread * math.cos(ang), rng.uniform(-0.15, 0.15), spread * math.sin(ang))
        xf = instance_transform(rng.normal(size=3), rng.uniform(0, 2 * math.pi),
                                rng.uniform(0.08, 0.2, 3), t)
        c = tuple(float(x) for x in rng.uniform(0.3, 0.9, 3))
        out.append(dict(vertices=v, indices=i, color=c, transform=xf, object_id=7,
                        instance_id=100 + k))
    return out


def instanced_pathtrace(count=6, level=3, vol_n=48, width=320, height=240, frames=1,
                        max_bounce=4, renderer="pathtrace"):
    """Instanced meshes around a Gaussian-blob volume under a directional
    light and ambient light: rotated, non-uniformly scaled instances of one
    icosphere (instanced_meshes), camera above the ring."""
    vox = _vrt.gen_blobs((vol_n, vol_n, vol_n), count=8, seed=3)
    sp = 0.6 / (vol_n - 1)
    pos = (0.9, 0.8, 1.1)
    g = max(1, (vol_n + 15) // 16)
    return Workload(
        name=f"{renderer} {count} instanced icospheres + blobs {vol_n}^3 {width}x{height}",
        voxels=vox, origin=(-0.3,) * 3, spacing=(sp,) * 3, tf=ramp_tf(256),
        value_range=(0.0, 1.0), unit_distance=0.05, grid_dims=(g, g, g),
        camera=dict(position=pos, direction=look_at(pos), up=(0, 1, 0),
                    fovy=math.radians(45.0), aspect=width / height),
        lights=[dict(type="directional", direction=(-0.3, -1.0, -0.4), color=(1, 1, 1),
                     intensity=2.0)],
        params=dict(background=(0.1, 0.1, 0.1, 1), ambient_color=(1, 1, 1),
                    ambient_radiance=0.3, occlusion_distance=1e20, ambient_samples=1,
                    max_bounce=max_bounce, pixel_samples=1, volume_sampling_rate=1.0),
        width=width, height=height, renderer=renderer, frames=frames,
        instances=instanced_meshes(count, level))


def _orbit_camera(seed, radius, height, width, aspect_h, fovy_deg=40.0):
    """Camera on a circle of `radius` around the y axis at `height`, at a
    seeded angle, looking at the origin."""
    u = _vrt.splitmix_uniform(seed, 1)[0]
    a = 2.0 * math.pi * u
    pos = (radius * math.cos(a), height, radius * math.sin(a))
    return dict(position=pos, direction=look_at(pos), up=(0, 1, 0),
                fovy=math.radians(fovy_deg), aspect=width / aspect_h)


def tets_pathtrace(n=70, width=3840, height=2160, frames=64, jitter=0.15, seed=5,
                   density_scale=40.0, albedo=0.8, max_bounce=8, grid=64, camera_seed=11,
                   threads=0):
    """Pathtrace an unstructured field: a jittered lattice of n^3 cubes over
    [-0.5, 0.5]^3 split into 6 n^3 tetrahedra (n = 70: 2,058,000) with
    Gaussian-blob vertex values (32 blobs); TF rgb = albedo, alpha = value,
    unitDistance = 1/density_scale; the reference's 64^3 majorant grid over
    the bounds; directional light + ambient; seeded orbit camera."""
    pos, val, idx = _vrt.gen_tet_lattice(n, jitter, 32, seed, threads)
    return Workload(
        name=f"pathtrace tets {len(idx)} {width}x{height}", voxels=None, origin=None,
        spacing=None, tf=albedo_tf(albedo, 256), value_range=(0.0, 1.0),
        unit_distance=1.0 / density_scale, grid_dims=(grid, grid, grid),
        camera=_orbit_camera(camera_seed, 1.9, 0.7, width, height),
        lights=[dict(type="directional", direction=(-0.4, -1.0, -0.3), color=(1, 1, 1),
                     intensity=1.0)],
        params=dict(background=(0, 0, 0, 1), ambient_color=(1, 1, 1), ambient_radiance=0.5,
                    occlusion_distance=1e20, ambient_samples=1, max_bounce=max_bounce,
                    pixel_samples=1),
        width=width, height=height, renderer="pathtrace", frames=frames,
        tets=dict(positions=pos, index=idx, vertex_data=val))


def tets_raycast(n=12, width=128, height=96, jitter=0.15, seed=5, threads=0):
    """Raycast the tetrahedral lattice field (ramp TF, step = cell size / 2)."""
    pos, val, idx = _vrt.gen_tet_lattice(n, jitter, 32, seed, threads)
    p = (0.0, 0.0, 2.5)
    return Workload(
        name=f"raycast tets {len(idx)} {width}x{height}", voxels=None, origin=None,
        spacing=None, tf=ramp_tf(256), value_range=(0.0, 1.0), unit_distance=1.0 / n,
        grid_dims=None,
        camera=dict(position=p, direction=look_at(p), up=(0, 1, 0),
                    fovy=math.radians(40.0), aspect=width / height),
        lights=[],
        params=dict(background=(0, 0, 0, 1), volume_sampling_rate=2.0, pixel_samples=1),
        width=width, height=height, renderer="raycast",
        tets=dict(positions=pos, index=idx, vertex_data=val))


def _grid_bounds(buf):
    """GridData::mWorldBBox of a NanoVDB buffer as float (lo xyz, hi xyz)."""
    return np.frombuffer(buf[560:608].tobytes(), np.float64).astype(np.float32)


def volume_bounds(wl):
    """Object-space bounds of the workload's volume."""
    if wl.nanovdb is not None:
        bb = _grid_bounds(wl.nanovdb)
        return bb[:3], bb[3:]
    if wl.tets is not None:
        p = wl.tets["positions"][wl.tets["index"].ravel()]
        return p.min(axis=0), p.max(axis=0)
    lo = np.array(wl.origin, np.float32)
    hi = lo + (np.array(wl.dims, np.float32) - 1) * np.array(wl.spacing, np.float32)
    return lo, hi


# ---------------------------------------------------------------------------
# GPU side
# ---------------------------------------------------------------------------
def light_desc(L, hdri=None):
    """`hdri`: the DeviceHdri of an "hdri" light (its "radiance" map)."""
    if L.get("type") == "hdri":
        return _vrt.hdri_light(hdri, L.get("up", (1, 0, 0)), L.get("direction", (0, 0, 1)),
                               L.get("scale", 1.0), L.get("visible", True))
    if L.get("type", "directional") == "quad":
        return _vrt.quad_light(L["position"], L["edge1"], L["edge2"], L["intensity"], L["color"],
                               L.get("visible", True))
    return _vrt.directional_light(L["direction"], L["intensity"], L["color"],
                                  L.get("visible", True))

class GpuScene:
    """The workload resident on one GPU through the vrt C ABI."""

    def __init__(self, ctx, wl: Workload, color_type=None, depth=False, accum64=False,
                 volume=None, params=None, bvh_builder="sah", gbuffer=False, taa=False):
        self.wl = wl
        self.ctx = ctx
        if volume is None and wl.nanovdb is not None:
            volume = _vrt.DeviceVolume.from_nanovdb(ctx, wl.nanovdb, wl.tf, wl.value_range,
                                                    wl.unit_distance, wl.grid_dims,
                                                    filter=wl.filter)
        if volume is None and wl.tets is not None:
            t = wl.tets
            field = _vrt.DeviceUnstructured(ctx, t["positions"], t["index"], t.get("cell_index"),
                                            t.get("cell_type"), t.get("vertex_data"),
                                            t.get("cell_data"))
            volume = _vrt.DeviceVolume.from_unstructured(ctx, field, wl.tf, wl.value_range,
                                                         wl.unit_distance, wl.grid_dims)
        if volume is None and wl.voxels is not None:
            volume = _vrt.DeviceVolume(ctx, wl.voxels, wl.origin, wl.spacing, wl.tf,
                                       wl.value_range, wl.unit_distance, wl.grid_dims)
        self.volume = volume
        self.hdris = [_vrt.DeviceHdri(ctx, L["radiance"]) if L.get("type") == "hdri" else None
                      for L in wl.lights]
        lights = [light_desc(L, h) for L, h in zip(wl.lights, self.hdris)]
        self.meshes = [(_vrt.DeviceMesh(ctx, m["vertices"], m.get("indices"), bvh_builder),
                        m["color"]) for m in (wl.meshes or [])]
        self.inst_meshes = [_vrt.DeviceMesh(ctx, m["vertices"], m.get("indices"), bvh_builder)
                            for m in (wl.instances or [])]
        self.tls = None
        if wl.instances:
            self.tls = _vrt.DeviceTls(ctx, [
                (mesh, m["transform"], m["color"], m.get("object_id", _vrt.NO_ID),
                 m.get("instance_id", _vrt.NO_ID))
                for mesh, m in zip(self.inst_meshes, wl.instances)])
        self.scene = _vrt.make_scene([volume] if volume is not None else [], lights, self.meshes,
                                     self.tls)
        self.camera = _vrt.camera(**wl.camera)
        p = dict(wl.params)
        p.update(params or {})
        self.params = _vrt.params(**p)
        self.frame = _vrt.DeviceFrame(ctx, wl.width, wl.height,
                                      color_type or _vrt.TYPE_FLOAT32_VEC4, depth, accum64,
                                      gbuffer=gbuffer, taa=taa)

    def render(self, accum_id=0):
        if self.wl.renderer == "raycast":
            _vrt.raycast(self.ctx, self.scene, self.camera, self.params, self.frame, accum_id)
        else:
            _vrt.pathtrace(self.ctx, self.scene, self.camera, self.params, self.frame, accum_id)

    def nrc_options(self, train=True, train_samples=65536, train_steps=4, query_vertex=2,
                    train_extra_vertices=0):
        """Cache options whose [0,1]^3 box is the volume's bounds."""
        lo, hi = volume_bounds(self.wl)
        return _vrt.nrc_render_options(lo, hi, train, train_samples, train_steps, query_vertex,
                                       train_extra_vertices)

    def render_nrc(self, nrc, opts, accum_id=0):
        _vrt.pathtrace_nrc(self.ctx, self.scene, self.camera, self.params, self.frame, nrc, opts,
                           accum_id)
