"""ctypes binding of the CPU oracle (oracle/vrt_oracle.h) -- test infrastructure.

Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only as the checker / the timed CPU restatement.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libvrt_oracle.so")
REF_NVDB_PATH = os.path.join(ROOT, "oracle", "_ref", "libnvdb_ref.so")

f3 = C.c_float * 3


class OrcLight(C.Structure):
    _fields_ = [("type", C.c_int), ("visible", C.c_int), ("dir", f3), ("pos", f3),
                ("e1", f3), ("e2", f3), ("color", f3), ("intensity", C.c_float),
                ("hdri_rgb", C.c_void_p), ("hdri_w", C.c_int), ("hdri_h", C.c_int),
                ("up", f3), ("scale", C.c_float)]


class OrcVolume(C.Structure):
    _fields_ = [("voxels", C.c_void_p), ("dims", C.c_int * 3), ("origin", f3), ("spacing", f3),
                ("tf_rgba", C.c_void_p), ("tf_size", C.c_int), ("value_lo", C.c_float),
                ("value_hi", C.c_float), ("unit_distance", C.c_float),
                ("majorants", C.c_void_p), ("grid_dims", C.c_int * 3),
                ("xfm", C.c_float * 12), ("field_type", C.c_int), ("nanovdb", C.c_void_p),
                ("filter", C.c_int), ("vertex_position", C.c_void_p),
                ("vertex_data", C.c_void_p), ("num_vertices", C.c_int64),
                ("index", C.c_void_p), ("num_indices", C.c_int64),
                ("cell_index", C.c_void_p), ("cell_type", C.c_void_p),
                ("cell_data", C.c_void_p), ("num_cells", C.c_int64),
                ("object_id", C.c_uint32), ("instance_id", C.c_uint32)]


class OrcTrimesh(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("indices", C.c_void_p), ("num_triangles", C.c_int),
                ("color", f3), ("opacity", C.c_float), ("object_id", C.c_uint32),
                ("instance_id", C.c_uint32)]


class OrcInstance(C.Structure):
    _fields_ = [("mesh", OrcTrimesh), ("affine_inv", C.c_float * 9), ("trans_inv", f3),
                ("normal_xfm", C.c_float * 9)]


class OrcScene(C.Structure):
    _fields_ = [("num_volumes", C.c_int), ("volumes", C.POINTER(OrcVolume)),
                ("num_lights", C.c_int), ("lights", C.POINTER(OrcLight)),
                ("num_meshes", C.c_int), ("meshes", C.POINTER(OrcTrimesh)),
                ("num_instances", C.c_int), ("instances", C.POINTER(OrcInstance))]


class OrcCamera(C.Structure):
    _fields_ = [("pos", f3), ("dir", f3), ("up", f3), ("fovy", C.c_float),
                ("aspect", C.c_float), ("region", C.c_float * 4), ("shutter", C.c_float * 2),
                ("aperture_radius", C.c_float), ("focus_distance", C.c_float)]


class OrcParams(C.Structure):
    _fields_ = [("bg", C.c_float * 4), ("ambient_color", f3), ("ambient_radiance", C.c_float),
                ("occlusion_distance", C.c_float), ("ambient_samples", C.c_int),
                ("max_bounce", C.c_int), ("pixel_samples", C.c_int),
                ("volume_sampling_rate", C.c_float), ("rng_offset", C.c_uint32),
                ("num_clip_planes", C.c_int), ("clip_planes", (C.c_float * 4) * 8),
                ("phase_g", C.c_float), ("gradient_shading", C.c_int),
                ("ratio_tracking", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run python oracle/build.py")
        L = C.CDLL(LIB_PATH)
        vp, ip = C.c_void_p, C.c_int
        L.orc_tea_seed.restype = C.c_uint32
        L.orc_tea_seed.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_lcg_stream.argtypes = [C.c_uint32, C.c_uint32, ip, vp]
        L.orc_tf_build.argtypes = [vp, ip, ip, vp, ip, vp, vp]
        L.orc_sample_field.argtypes = [C.POINTER(OrcVolume), ip, vp, vp]
        L.orc_post_classify.argtypes = [C.POINTER(OrcVolume), ip, vp, vp]
        L.orc_build_majorants.argtypes = [C.POINTER(OrcVolume), ip, ip, ip, vp]
        L.orc_build_majorants_reference.argtypes = [C.POINTER(OrcVolume), ip, ip, ip, vp]
        L.orc_dda_cells.argtypes = [vp, vp, C.c_float, C.c_float, vp, vp, vp, ip, vp, vp, vp]
        L.orc_delta_track.argtypes = [C.POINTER(OrcVolume), vp, vp, C.c_float, C.c_float,
                                      C.POINTER(C.c_uint32), C.POINTER(C.c_float), ip]
        L.orc_ratio_track.restype = C.c_float
        L.orc_ratio_track.argtypes = [C.POINTER(OrcVolume), vp, vp, C.c_float, C.c_float,
                                      C.POINTER(C.c_uint32), ip]
        L.orc_raycast.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcCamera),
                                  C.POINTER(OrcParams), ip, ip, vp, ip]
        L.orc_pathtrace_frame.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcCamera),
                                          C.POINTER(OrcParams), ip, ip, ip, ip, ip, ip, ip, ip,
                                          vp, ip, C.POINTER(C.c_uint64)]
        L.orc_pathtrace_gbuffer.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcCamera),
                                            C.POINTER(OrcParams), ip, ip, ip, vp, vp, vp, vp,
                                            vp, vp, ip]
        L.orc_pathtrace_taa.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcCamera),
                                         C.POINTER(OrcCamera), C.POINTER(OrcParams), ip, ip, ip,
                                         C.c_float, vp, vp, vp, vp, ip]
        L.orc_raycast_gbuffer.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcCamera),
                                          C.POINTER(OrcParams), ip, ip, vp, vp, vp, vp, vp,
                                          vp, ip]
        L.orc_pathtrace_accumulate.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcCamera),
                                               C.POINTER(OrcParams), ip, ip, ip, ip, ip, ip,
                                               vp, ip]
        L.orc_cvt_uint32.restype = C.c_uint32
        L.orc_cvt_uint32.argtypes = [vp]
        L.orc_cvt_uint32_srgb.restype = C.c_uint32
        L.orc_cvt_uint32_srgb.argtypes = [vp]
        L.orc_camera_ray.argtypes = [C.POINTER(OrcCamera), C.c_float, C.c_float, ip, ip, vp, vp]
        L.orc_camera_lens_ray.argtypes = [C.POINTER(OrcCamera), C.c_float, C.c_float, ip, ip,
                                          C.c_float, C.c_float, vp, vp]
        L.orc_mesh_intersect.argtypes = [C.POINTER(OrcTrimesh), ip, vp, ip, ip, vp, vp]
        L.orc_scene_intersect_surfaces.argtypes = [C.POINTER(OrcScene), ip, vp, ip, vp, vp, vp]
        L.orc_instance_transforms.argtypes = [vp, vp, vp, vp]
        L.orc_quad_light_sample.argtypes = [C.POINTER(OrcLight), vp, ip, C.c_uint32, vp, vp]
        L.orc_unstructured_info.argtypes = [C.POINTER(OrcVolume), vp, vp, vp, vp]
        L.orc_sample_field_brute.argtypes = [C.POINTER(OrcVolume), ip, vp, vp]
        L.orc_hg_sample.argtypes = [C.c_float, vp, ip, C.c_uint32, vp]
        L.orc_uniform_sample_sphere.argtypes = [ip, C.c_uint32, vp]
        L.orc_pow_log2_table.argtypes = [vp]
        L.orc_set_roulette.argtypes = [ip]
        L.orc_hg_4pi.restype = C.c_double
        L.orc_hg_4pi.argtypes = [C.c_double, C.c_double]
        L.orc_hdri_cdf.argtypes = [C.POINTER(OrcLight), vp, vp]
        L.orc_hdri_sample.argtypes = [C.POINTER(OrcLight), ip, C.c_uint32, vp, vp, vp]
        L.orc_hdri_eval.argtypes = [C.POINTER(OrcLight), ip, vp, vp, vp, vp]
        L.orc_nvdb_get.argtypes = [vp, ip, vp, vp]
        L.orc_nvdb_sample.argtypes = [vp, ip, vp, ip, vp]
        L.orc_nvdb_world_bbox.argtypes = [vp, vp]
        _lib = L
    return _lib


_ref_nvdb = None


def ref_nvdb():
    """The reference's vendored NanoVDB compiled into oracle/_ref (None when it
    has not been built: the reference tree is absent)."""
    global _ref_nvdb
    if _ref_nvdb is None and os.path.exists(REF_NVDB_PATH):
        R = C.CDLL(REF_NVDB_PATH)
        vp, ip = C.c_void_p, C.c_int
        R.ref_nvdb_check.argtypes = [vp, C.c_char_p, ip]
        R.ref_nvdb_get.argtypes = [vp, ip, vp, vp]
        R.ref_nvdb_sample.argtypes = [vp, ip, vp, ip, vp]
        R.ref_nvdb_world_bbox.argtypes = [vp, vp]
        R.ref_nvdb_make_tiled.argtypes = [C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        R.ref_nvdb_free.argtypes = [vp]
        _ref_nvdb = R
    return _ref_nvdb


def nvdb_get(buf, ijk):
    ijk = np.ascontiguousarray(ijk, np.int32)
    out = np.empty(len(ijk), np.float32)
    assert lib().orc_nvdb_get(buf.ctypes.data, len(ijk), ijk.ctypes.data, out.ctypes.data) == 0
    return out


def nvdb_sample(buf, xyz, order=1):
    xyz = np.ascontiguousarray(xyz, np.float32)
    out = np.empty(len(xyz), np.float32)
    assert lib().orc_nvdb_sample(buf.ctypes.data, len(xyz), xyz.ctypes.data, order,
                                 out.ctypes.data) == 0
    return out


def nvdb_world_bbox(buf):
    out = np.empty(6, np.float64)
    assert lib().orc_nvdb_world_bbox(buf.ctypes.data, out.ctypes.data) == 0
    return out


def lcg_stream(s0, s1, n):
    out = np.empty(n, np.float32)
    lib().orc_lcg_stream(s0, s1, n, out.ctypes.data)
    return out


def tf_build(colors=None, opacities=None, uniform=(1, 1, 1, 1)):
    uni = np.asarray(uniform, np.float32)
    cptr, nc, nch = None, 0, 4
    if colors is not None:
        colors = np.ascontiguousarray(colors, np.float32)
        nc, nch = colors.shape
        cptr = colors.ctypes.data
    optr, no = None, 0
    if opacities is not None:
        opacities = np.ascontiguousarray(opacities, np.float32)
        no = opacities.shape[0]
        optr = opacities.ctypes.data
    n = lib().orc_tf_build(cptr, nc, nch, optr, no, uni.ctypes.data, None)
    out = np.empty((n, 4), np.float32)
    lib().orc_tf_build(cptr, nc, nch, optr, no, uni.ctypes.data, out.ctypes.data)
    return out


class OracleVolume:
    """Keeps numpy arrays alive while the struct points at them."""

    def __init__(self, voxels, origin, spacing, tf, value_range, unit_distance,
                 majorants=None, grid_dims=None, xfm=None):
        self.voxels = np.ascontiguousarray(voxels, np.float32)
        self.tf = np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
        nz, ny, nx = self.voxels.shape
        v = OrcVolume()
        v.object_id = v.instance_id = 0xFFFFFFFF
        v.voxels = self.voxels.ctypes.data
        v.dims = (C.c_int * 3)(nx, ny, nz)
        v.origin = f3(*origin)
        v.spacing = f3(*spacing)
        v.tf_rgba = self.tf.ctypes.data
        v.tf_size = self.tf.shape[0]
        v.value_lo, v.value_hi = value_range
        v.unit_distance = unit_distance
        self.majorants = None
        if majorants is not None:
            self.majorants = np.ascontiguousarray(majorants, np.float32)
            v.majorants = self.majorants.ctypes.data
            v.grid_dims = (C.c_int * 3)(*grid_dims)
        x = np.eye(3, 4, dtype=np.float32) if xfm is None else np.asarray(xfm, np.float32)
        v.xfm = (C.c_float * 12)(*x.ravel())
        self.s = v

    @classmethod
    def from_nanovdb(cls, buf, tf, value_range, unit_distance, filter="linear", xfm=None):
        self = cls.__new__(cls)
        self.voxels = None
        self.buf = buf
        self.tf = np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
        v = OrcVolume()
        v.object_id = v.instance_id = 0xFFFFFFFF
        v.field_type = 1
        v.nanovdb = buf.ctypes.data
        v.filter = {"linear": 0, "nearest": 1}[filter]
        v.tf_rgba = self.tf.ctypes.data
        v.tf_size = self.tf.shape[0]
        v.value_lo, v.value_hi = value_range
        v.unit_distance = unit_distance
        self.majorants = None
        x = np.eye(3, 4, dtype=np.float32) if xfm is None else np.asarray(xfm, np.float32)
        v.xfm = (C.c_float * 12)(*x.ravel())
        self.s = v
        return self

    @classmethod
    def from_unstructured(cls, positions, index, tf, value_range, unit_distance,
                          vertex_data=None, cell_index=None, cell_type=None, cell_data=None,
                          xfm=None):
        """An unstructured field from its ANARI arrays (tetrahedra only when
        cell_index / cell_type are omitted)."""
        self = cls.__new__(cls)
        self.voxels = None
        self.pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(index, np.uint32).ravel()
        if cell_index is None:
            cell_index = np.arange(len(self.idx) // 4, dtype=np.uint32) * 4
            cell_type = np.full(len(self.idx) // 4, 10, np.uint8)
        self.ci = np.ascontiguousarray(cell_index, np.uint32)
        self.ct = np.ascontiguousarray(cell_type, np.uint8)
        self.vd = None if vertex_data is None else np.ascontiguousarray(vertex_data, np.float32)
        self.cd = None if cell_data is None else np.ascontiguousarray(cell_data, np.float32)
        self.tf = np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
        v = OrcVolume()
        v.object_id = v.instance_id = 0xFFFFFFFF
        v.field_type = 2
        v.vertex_position = self.pos.ctypes.data
        v.vertex_data = None if self.vd is None else self.vd.ctypes.data
        v.num_vertices = len(self.pos)
        v.index = self.idx.ctypes.data
        v.num_indices = len(self.idx)
        v.cell_index = self.ci.ctypes.data
        v.cell_type = self.ct.ctypes.data
        v.cell_data = None if self.cd is None else self.cd.ctypes.data
        v.num_cells = len(self.ci)
        v.tf_rgba = self.tf.ctypes.data
        v.tf_size = self.tf.shape[0]
        v.value_lo, v.value_hi = value_range
        v.unit_distance = unit_distance
        self.majorants = None
        x = np.eye(3, 4, dtype=np.float32) if xfm is None else np.asarray(xfm, np.float32)
        v.xfm = (C.c_float * 12)(*x.ravel())
        self.s = v
        return self

    def unstructured_info(self):
        """(bounds (6,), cell_size, cells kept, cells skipped), or None when the
        field holds non-tetrahedral cells."""
        b = np.empty(6, np.float32)
        cs = C.c_float()
        kept, skipped = C.c_int64(), C.c_int64()
        if lib().orc_unstructured_info(C.byref(self.s), b.ctypes.data, C.byref(cs),
                                       C.byref(kept), C.byref(skipped)):
            return None
        return b, cs.value, kept.value, skipped.value

    def sample_brute(self, p_voxel):
        p = np.ascontiguousarray(p_voxel, np.float32).reshape(-1, 3)
        out = np.empty(p.shape[0], np.float32)
        lib().orc_sample_field_brute(C.byref(self.s), p.shape[0], p.ctypes.data,
                                     out.ctypes.data)
        return out

    def build_majorants(self, grid_dims):
        gx, gy, gz = grid_dims
        out = np.empty((gz, gy, gx), np.float32)
        lib().orc_build_majorants(C.byref(self.s), gx, gy, gz, out.ctypes.data)
        return out

    def attach_majorants(self, grid_dims):
        m = self.build_majorants(grid_dims)
        self.majorants = m
        self.s.majorants = m.ctypes.data
        self.s.grid_dims = (C.c_int * 3)(*grid_dims)
        return m

    def sample(self, p_voxel):
        p = np.ascontiguousarray(p_voxel, np.float32).reshape(-1, 3)
        out = np.empty(p.shape[0], np.float32)
        lib().orc_sample_field(C.byref(self.s), p.shape[0], p.ctypes.data, out.ctypes.data)
        return out

    def classify(self, values):
        v = np.ascontiguousarray(values, np.float32).ravel()
        out = np.empty((v.shape[0], 4), np.float32)
        lib().orc_post_classify(C.byref(self.s), v.shape[0], v.ctypes.data, out.ctypes.data)
        return out


class OracleMesh:
    """A matte triangle mesh (float3 vertices, uint3 indices or a soup)."""

    def __init__(self, vertices, indices=None, color=(0.7, 0.7, 0.7)):
        self.v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.i = None if indices is None else np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        m = OrcTrimesh()
        m.object_id = m.instance_id = 0xFFFFFFFF
        m.vertices = self.v.ctypes.data
        m.indices = None if self.i is None else self.i.ctypes.data
        m.num_triangles = len(self.v) // 3 if self.i is None else len(self.i)
        m.color = f3(*color)
        m.opacity = 1.0
        self.s = m


def instance_transforms(transform):
    """The oracle's (affine_inv, trans_inv, normal_xfm) of a row-major 3x4
    [L | t]; None for a singular L."""
    x = np.ascontiguousarray(transform, np.float32).reshape(12)
    a, t, n = np.empty(9, np.float32), np.empty(3, np.float32), np.empty(9, np.float32)
    if lib().orc_instance_transforms(x.ctypes.data, a.ctypes.data, t.ctypes.data,
                                     n.ctypes.data) != 0:
        return None
    return a.reshape(3, 3), t, n.reshape(3, 3)


class OracleInstance:
    """An OracleMesh placed by a row-major 3x4 transform (its inverse and
    normal transform from instance_transforms)."""

    def __init__(self, mesh, transform, object_id=0xFFFFFFFF, instance_id=0xFFFFFFFF):
        self.mesh = mesh
        a, t, n = instance_transforms(transform)
        s = OrcInstance()
        s.mesh = mesh.s
        s.mesh.object_id = object_id
        s.mesh.instance_id = instance_id
        s.affine_inv = (C.c_float * 9)(*a.ravel())
        s.trans_inv = f3(*t)
        s.normal_xfm = (C.c_float * 9)(*n.ravel())
        self.s = s


class OracleScene:
    def __init__(self, volumes, lights=(), meshes=(), instances=()):
        self.volumes = list(volumes)
        self.meshes = list(meshes)
        self.instances = list(instances)
        self.varr = (OrcVolume * max(1, len(self.volumes)))(*[v.s for v in self.volumes])
        self.marr = (OrcTrimesh * max(1, len(self.meshes)))(*[m.s for m in self.meshes])
        self.iarr = (OrcInstance * max(1, len(self.instances)))(*[i.s for i in self.instances])
        self.larr = (OrcLight * max(1, len(lights)))()
        for i, L in enumerate(lights):
            self.larr[i] = L
        s = OrcScene()
        s.num_volumes = len(self.volumes)
        s.volumes = C.cast(self.varr, C.POINTER(OrcVolume))
        s.num_lights = len(lights)
        s.lights = C.cast(self.larr, C.POINTER(OrcLight))
        s.num_meshes = len(self.meshes)
        s.meshes = C.cast(self.marr, C.POINTER(OrcTrimesh))
        s.num_instances = len(self.instances)
        s.instances = C.cast(self.iarr, C.POINTER(OrcInstance))
        self.s = s


def light_quad(position, edge1, edge2, intensity=1.0, color=(1, 1, 1), visible=True):
    L = OrcLight()
    L.type = 1
    L.visible = int(visible)
    L.pos = f3(*position)
    L.e1 = f3(*edge1)
    L.e2 = f3(*edge2)
    L.color = f3(*color)
    L.intensity = max(0.0, intensity)
    return L


_HDRI_KEEP = []


def light_hdri(radiance, up=(1, 0, 0), direction=(0, 0, 1), scale=1.0, visible=True):
    """An "hdri" light over the (h, w, 3) float32 map (kept alive here)."""
    rgb = np.ascontiguousarray(radiance, np.float32)
    _HDRI_KEEP.append(rgb)
    L = OrcLight()
    L.type = 2
    L.visible = int(visible)
    L.hdri_rgb = rgb.ctypes.data
    L.hdri_h, L.hdri_w = rgb.shape[:2]
    L.up = f3(*up)
    L.dir = f3(*direction)
    L.scale = scale
    return L


def light_directional(direction, intensity=1.0, color=(1, 1, 1), visible=True):
    L = OrcLight()
    L.type = 0
    L.visible = int(visible)
    L.dir = f3(*direction)
    L.color = f3(*color)
    L.intensity = intensity
    return L


def camera(position, direction, up, fovy, aspect, region=(0, 0, 1, 1), shutter=(0.5, 0.5),
           aperture_radius=0.0, focus_distance=1.0):
    c = OrcCamera()
    c.pos = f3(*position)
    c.dir = f3(*direction)
    c.up = f3(*up)
    c.fovy = fovy
    c.aspect = aspect
    c.region = (C.c_float * 4)(*region)
    c.shutter = (C.c_float * 2)(*shutter)
    c.aperture_radius = aperture_radius
    c.focus_distance = focus_distance
    return c


def params(background=(0, 0, 0, 1), ambient_color=(1, 1, 1), ambient_radiance=0.2,
           occlusion_distance=1e20, ambient_samples=1, max_bounce=7, pixel_samples=1,
           volume_sampling_rate=0.5, rng_offset=0, clip_planes=(), phase_g=0.0,
           gradient_shading=False, taa=False, taa_alpha=0.3, ratio_tracking=False):
    # taa / taa_alpha: the oracle's TAA frames take them as arguments
    # (OracleWorkload.pathtrace_taa)
    p = OrcParams()
    p.ratio_tracking = int(bool(ratio_tracking))
    p.phase_g = phase_g
    p.gradient_shading = int(bool(gradient_shading))
    p.num_clip_planes = len(clip_planes)
    for i, c in enumerate(clip_planes):
        p.clip_planes[i] = (C.c_float * 4)(*c)
    p.bg = (C.c_float * 4)(*background)
    p.ambient_color = f3(*ambient_color)
    p.ambient_radiance = ambient_radiance
    p.occlusion_distance = occlusion_distance
    p.ambient_samples = ambient_samples
    p.max_bounce = max_bounce
    p.pixel_samples = pixel_samples
    p.volume_sampling_rate = volume_sampling_rate
    p.rng_offset = rng_offset
    return p


class OracleWorkload:
    """A workloads.Workload restated for the oracle."""

    def __init__(self, wl):
        self.wl = wl
        self.vol = None
        if wl.nanovdb is not None:
            self.vol = OracleVolume.from_nanovdb(wl.nanovdb, wl.tf, wl.value_range,
                                                 wl.unit_distance, wl.filter)
        elif getattr(wl, "tets", None) is not None:
            t = wl.tets
            self.vol = OracleVolume.from_unstructured(
                t["positions"], t["index"], wl.tf, wl.value_range, wl.unit_distance,
                t.get("vertex_data"), t.get("cell_index"), t.get("cell_type"), t.get("cell_data"))
        elif wl.voxels is not None:
            self.vol = OracleVolume(wl.voxels, wl.origin, wl.spacing, wl.tf, wl.value_range,
                                    wl.unit_distance)
        if self.vol is not None and wl.grid_dims is not None:
            self.vol.attach_majorants(wl.grid_dims)
        lights = []
        for L in wl.lights:
            if L.get("type") == "hdri":
                lights.append(light_hdri(L["radiance"], L.get("up", (1, 0, 0)),
                                         L.get("direction", (0, 0, 1)), L.get("scale", 1.0),
                                         L.get("visible", True)))
            elif L.get("type", "directional") == "quad":
                lights.append(light_quad(L["position"], L["edge1"], L["edge2"], L["intensity"],
                                         L["color"], L.get("visible", True)))
            else:
                lights.append(light_directional(L["direction"], L["intensity"], L["color"],
                                                L.get("visible", True)))
        self.meshes = [OracleMesh(m["vertices"], m.get("indices"), m["color"])
                       for m in (wl.meshes or [])]
        self.instances = [
            OracleInstance(OracleMesh(m["vertices"], m.get("indices"), m["color"]),
                           m["transform"], m.get("object_id", 0xFFFFFFFF),
                           m.get("instance_id", 0xFFFFFFFF))
            for m in (getattr(wl, "instances", None) or [])]
        self.scene = OracleScene([self.vol] if self.vol is not None else [], lights, self.meshes,
                                 self.instances)
        self.cam = camera(**wl.camera)
        self.params = params(**wl.params)

    def raycast(self, threads=0):
        out = np.empty((self.wl.height, self.wl.width, 4), np.float32)
        lib().orc_raycast(C.byref(self.scene.s), C.byref(self.cam), C.byref(self.params),
                          self.wl.width, self.wl.height, out.ctypes.data, threads)
        return out

    def pathtrace_frame(self, accum_id=0, fp64=False, stride=1, y0=0, y1=0, threads=0):
        W, H = self.wl.width, self.wl.height
        if stride > 1:
            shape = ((H + stride - 1) // stride, (W + stride - 1) // stride, 4)
        else:
            shape = (H, W, 4)
        out = np.zeros(shape, np.float32)
        lookups = C.c_uint64(0)
        lib().orc_pathtrace_frame(C.byref(self.scene.s), C.byref(self.cam), C.byref(self.params),
                                  W, H, accum_id, 0, int(fp64), stride, y0, y1, out.ctypes.data,
                                  threads, C.byref(lookups))
        return out, lookups.value

    def gbuffer(self, accum_ids=(0,), threads=0):
        """The frame's g-buffer after the frames `accum_ids` (from an
        accumulation reset: depth 1e31): dict of depth (h, w), normal and
        albedo (h, w, 3), primitive_id / object_id / instance_id (h, w)
        uint32."""
        W, H = self.wl.width, self.wl.height
        g = {"depth": np.full((H, W), 1e31, np.float32),
             "normal": np.zeros((H, W, 3), np.float32), "albedo": np.zeros((H, W, 3), np.float32),
             "primitive_id": np.zeros((H, W), np.uint32), "object_id": np.zeros((H, W), np.uint32),
             "instance_id": np.zeros((H, W), np.uint32)}
        ptrs = [g[k].ctypes.data for k in ("depth", "normal", "albedo", "primitive_id",
                                           "object_id", "instance_id")]
        for a in accum_ids:
            if self.wl.renderer == "raycast":
                lib().orc_raycast_gbuffer(C.byref(self.scene.s), C.byref(self.cam),
                                          C.byref(self.params), W, H, *ptrs, threads)
            else:
                lib().orc_pathtrace_gbuffer(C.byref(self.scene.s), C.byref(self.cam),
                                            C.byref(self.params), W, H, a, *ptrs, threads)
        return g

    def taa_state(self):
        """Zeroed TAA history (colour float4, albedo float3), as the reference
        allocates it."""
        W, H = self.wl.width, self.wl.height
        return {"rgba": np.zeros((H, W, 4), np.float32), "albedo": np.zeros((H, W, 3), np.float32)}

    def pathtrace_taa(self, accum_id, state, alpha=0.3, cam=None, prev_cam=None, threads=0):
        """One TAA frame against `state` (updated in place): the resolved
        image and the motion vectors (h, w, 2)."""
        W, H = self.wl.width, self.wl.height
        out = np.empty((H, W, 4), np.float32)
        mv = np.empty((H, W, 2), np.float32)
        lib().orc_pathtrace_taa(C.byref(self.scene.s), C.byref(cam or self.cam),
                                C.byref(prev_cam) if prev_cam is not None else None,
                                C.byref(self.params), W, H, accum_id, alpha,
                                state["rgba"].ctypes.data, state["albedo"].ctypes.data,
                                out.ctypes.data, mv.ctypes.data, threads)
        return out, mv

    def pathtrace_accumulate(self, frames, first_accum_id=0, fp64=False, stride=1, threads=0):
        W, H = self.wl.width, self.wl.height
        if stride > 1:
            shape = ((H + stride - 1) // stride, (W + stride - 1) // stride, 4)
        else:
            shape = (H, W, 4)
        out = np.zeros(shape, np.float64)
        lib().orc_pathtrace_accumulate(C.byref(self.scene.s), C.byref(self.cam),
                                       C.byref(self.params), W, H, first_accum_id, frames,
                                       int(fp64), stride, out.ctypes.data, threads)
        return out


# ---------------------------------------------------------------------------
# neural radiance cache MLP (oracle/nrc_oracle.cpp)
# ---------------------------------------------------------------------------
def _nrc_lib():
    L = lib()
    if not getattr(L, "_nrc_ready", False):
        L.orc_nrc_param_count.restype = C.c_int
        L.orc_nrc_encode.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_nrc_forward.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_nrc_gradients.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
        L.orc_nrc_gradients.restype = C.c_double
        L._nrc_ready = True
    return L


def nrc_forward(params, raw, precision):
    L = _nrc_lib()
    p = np.ascontiguousarray(params, np.float32)
    r = np.ascontiguousarray(raw, np.float32).reshape(-1, 6)
    out = np.empty((r.shape[0], 3), np.float64)
    L.orc_nrc_forward(p.ctypes.data, precision, r.shape[0], r.ctypes.data, out.ctypes.data)
    return out


def nrc_gradients(params, raw, target, precision):
    L = _nrc_lib()
    p = np.ascontiguousarray(params, np.float32)
    r = np.ascontiguousarray(raw, np.float32).reshape(-1, 6)
    t = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
    g = np.empty(L.orc_nrc_param_count(), np.float64)
    loss = L.orc_nrc_gradients(p.ctypes.data, precision, r.shape[0], r.ctypes.data,
                               t.ctypes.data, g.ctypes.data)
    return g, loss


def nrc_encode(raw):
    L = _nrc_lib()
    r = np.ascontiguousarray(raw, np.float32)
    out = np.empty(128, np.float64)
    L.orc_nrc_encode(r.ctypes.data, out.ctypes.data)
    return out


def scene_intersect_surfaces(scene, rays, any_hit=False):
    """(t, prim, surface) of rays (n, 8) over an OracleScene's meshes and
    instances (float arithmetic)."""
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    t = np.empty(len(r), np.float32)
    p = np.empty(len(r), np.int32)
    sf = np.empty(len(r), np.int32)
    lib().orc_scene_intersect_surfaces(C.byref(scene.s), len(r), r.ctypes.data, int(any_hit),
                                       t.ctypes.data, p.ctypes.data, sf.ctypes.data)
    return t, p, sf


def mesh_intersect(mesh, rays, any_hit=False, brute=False):
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    t = np.empty(len(r), np.float32)
    p = np.empty(len(r), np.int32)
    lib().orc_mesh_intersect(C.byref(mesh.s), len(r), r.ctypes.data, int(any_hit), int(brute),
                             t.ctypes.data, p.ctypes.data)
    return t, p


def quad_light_sample(light, ref, n, seed=1):
    ref = np.ascontiguousarray(ref, np.float32)
    pos = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    lib().orc_quad_light_sample(C.byref(light), ref.ctypes.data, n, seed, pos.ctypes.data,
                                pdf.ctypes.data)
    return pos, pdf


# Share of pixels whose GPU sample equals the oracle's to 1e-4 relative.  With
# the same RNG streams every path takes the same branches; a pixel differs
# only where float rounding flips a branch (a roulette or collision draw that
# lands within rounding of its threshold), which stays far below 2 %.
STREAM_AGREEMENT_MIN = 0.98


def stream_agreement(gpu, orc):
    close = np.all(np.abs(gpu - orc) <= 1e-4 * np.maximum(1, np.abs(orc)), axis=-1)
    return float(close.mean())
