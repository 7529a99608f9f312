"""ctypes binding of libvrt.so (include/vrt.h) and of the scene generators.

The binding is thin: structs mirror the C declarations field by field, and
every call checks the returned status and raises VrtError with the library's
message.  Loading fails loudly when the HIP library has not been built --
there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(PKG, "_build")
LIB_PATH = os.path.join(BUILD, "libvrt.so")
GEN_PATH = os.path.join(BUILD, "libvrt_scenegen.so")

FIELD_STRUCTURED_REGULAR = 0
FIELD_NANOVDB = 1
FIELD_UNSTRUCTURED = 2
LIGHT_DIRECTIONAL = 0
LIGHT_QUAD = 1
LIGHT_HDRI = 2
MAX_VOLUMES = 4
MAX_LIGHTS = 8
TYPE_FLOAT32 = 1068
TYPE_FLOAT32_VEC4 = 1071
TYPE_UFIXED8_VEC4 = 1039
TYPE_UFIXED8_RGBA_SRGB = 2003


class VrtError(RuntimeError):
    pass


f3 = C.c_float * 3
f4 = C.c_float * 4
i3 = C.c_int * 3


class Volume(C.Structure):
    _fields_ = [
        ("field_type", C.c_int), ("dims", i3), ("bounds_lo", f3), ("bounds_hi", f3),
        ("cell_size", C.c_float), ("bricks", C.c_void_p), ("nanovdb", C.c_void_p),
        ("filter", C.c_int), ("unstructured", C.c_void_p),
        ("tf_rgba", C.c_void_p), ("tf_size", C.c_int), ("value_lo", C.c_float),
        ("value_hi", C.c_float), ("unit_distance", C.c_float), ("majorants", C.c_void_p),
        ("grid_dims", i3), ("majorant_bytes", C.c_void_p), ("majorant_step", C.c_void_p),
        ("affine_inv", C.c_float * 9), ("trans_inv", f3),
        ("object_id", C.c_uint32), ("instance_id", C.c_uint32),
    ]


class Light(C.Structure):
    _fields_ = [
        ("type", C.c_int), ("visible", C.c_int), ("direction", f3), ("position", f3),
        ("edge1", f3), ("edge2", f3), ("color", f3), ("intensity", C.c_float),
        ("hdri", C.c_void_p), ("up", f3), ("scale", C.c_float),
    ]


class Surface(C.Structure):
    _fields_ = [("bvh", C.c_void_p), ("color", f3), ("object_id", C.c_uint32),
                ("instance_id", C.c_uint32)]


NO_ID = 0xFFFFFFFF  # an unset ANARI "id"


MAX_SURFACES = 4


class Instance(C.Structure):
    _fields_ = [("bvh", C.c_void_p), ("transform", C.c_float * 12), ("color", f3),
                ("object_id", C.c_uint32), ("instance_id", C.c_uint32)]


class Scene(C.Structure):
    _fields_ = [
        ("num_volumes", C.c_int), ("volumes", Volume * MAX_VOLUMES),
        ("num_lights", C.c_int), ("lights", Light * MAX_LIGHTS),
        ("num_surfaces", C.c_int), ("surfaces", Surface * MAX_SURFACES),
        ("tls", C.c_void_p),
    ]


class BvhInfo(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("num_triangles", C.c_uint32), ("depth", C.c_uint32),
                ("bounds_lo", f3), ("bounds_hi", f3), ("build_ms", C.c_double),
                ("device_bytes", C.c_uint64), ("stack_bound", C.c_uint32)]


class TlsInfo(C.Structure):
    _fields_ = [("num_instances", C.c_uint32), ("num_nodes", C.c_uint32),
                ("stack_bound", C.c_uint32), ("bounds_lo", f3), ("bounds_hi", f3),
                ("build_ms", C.c_double)]


class Camera(C.Structure):
    _fields_ = [
        ("position", f3), ("direction", f3), ("up", f3), ("fovy", C.c_float),
        ("aspect", C.c_float), ("image_region", f4), ("shutter", C.c_float * 2),
        ("aperture_radius", C.c_float), ("focus_distance", C.c_float),
    ]


class Params(C.Structure):
    _fields_ = [
        ("background", f4), ("ambient_color", f3), ("ambient_radiance", C.c_float),
        ("occlusion_distance", C.c_float), ("ambient_samples", C.c_int),
        ("max_bounce", C.c_int), ("pixel_samples", C.c_int),
        ("volume_sampling_rate", C.c_float), ("rng_offset", C.c_uint32),
        ("num_clip_planes", C.c_int), ("clip_planes", (C.c_float * 4) * 8),
        ("phase_g", C.c_float), ("gradient_shading", C.c_int),
        ("taa", C.c_int), ("taa_alpha", C.c_float), ("has_prev_camera", C.c_int),
        ("prev_camera", Camera), ("ratio_tracking", C.c_int), ("pass_schedule", C.c_int),
    ]


class Frame(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("color_type", C.c_int),
        ("color", C.c_void_p), ("accum", C.c_void_p), ("depth", C.c_void_p),
        ("accum64", C.c_void_p), ("counters", C.c_void_p),
        ("normal", C.c_void_p), ("albedo", C.c_void_p), ("primitive_id", C.c_void_p),
        ("object_id", C.c_void_p), ("instance_id", C.c_void_p),
        ("motion_vec", C.c_void_p), ("taa_curr", C.c_void_p), ("taa_prev", C.c_void_p),
        ("taa_curr_albedo", C.c_void_p), ("taa_prev_albedo", C.c_void_p),
    ]


class NrcConfig(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("learning_rate", C.c_float), ("beta1", C.c_float),
        ("beta2", C.c_float), ("epsilon", C.c_float), ("max_batch", C.c_int),
        ("seed", C.c_uint64),
    ]


class NrcRender(C.Structure):
    _fields_ = [
        ("train", C.c_int), ("train_samples", C.c_int), ("train_steps", C.c_int),
        ("bounds_lo", f3), ("bounds_hi", f3), ("query_vertex", C.c_int),
        ("train_extra_vertices", C.c_int),
    ]


NRC_FP16 = 0
NRC_BF16 = 1

_lib = None
_gen = None


def lib():
    """The loaded libvrt.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VrtError(f"{LIB_PATH} is missing: run python anari-visionaray_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, sz, ip = C.c_void_p, C.c_size_t, C.c_int
        L.vrt_last_error.restype = C.c_char_p
        L.vrt_context_create.argtypes = [ip, C.POINTER(vp)]
        L.vrt_context_destroy.argtypes = [vp]
        L.vrt_context_stream.argtypes = [vp, C.POINTER(vp)]
        L.vrt_synchronize.argtypes = [vp]
        L.vrt_alloc.argtypes = [vp, sz, C.POINTER(vp)]
        L.vrt_free.argtypes = [vp, vp]
        L.vrt_upload.argtypes = [vp, vp, vp, sz]
        L.vrt_download.argtypes = [vp, vp, vp, sz]
        L.vrt_memset.argtypes = [vp, vp, ip, sz]
        L.vrt_device_copy.argtypes = [vp, vp, vp, sz]
        L.vrt_timer_start.argtypes = [vp]
        L.vrt_timer_stop.argtypes = [vp, C.POINTER(C.c_float)]
        L.vrt_structured_brick_bytes.argtypes = [C.POINTER(C.c_int)]
        L.vrt_structured_brick_bytes.restype = sz
        L.vrt_structured_bricks.argtypes = [vp, vp, C.POINTER(C.c_int), vp]
        L.vrt_build_value_ranges.argtypes = [vp, C.POINTER(Volume), vp]
        L.vrt_compute_majorants.argtypes = [vp, C.POINTER(Volume), vp, vp, vp, vp]
        L.vrt_frame_reset.argtypes = [vp, C.POINTER(Frame)]
        L.vrt_nrc_param_count.restype = ip
        L.vrt_nrc_create.argtypes = [vp, C.POINTER(NrcConfig), C.POINTER(vp)]
        L.vrt_nrc_destroy.argtypes = [vp]
        L.vrt_nrc_get_params.argtypes = [vp, vp]
        L.vrt_nrc_set_params.argtypes = [vp, vp]
        L.vrt_nrc_infer.argtypes = [vp, ip, vp, vp]
        L.vrt_nrc_gradients.argtypes = [vp, ip, vp, vp, vp, C.POINTER(C.c_float)]
        L.vrt_nrc_train_step.argtypes = [vp, ip, vp, vp, C.POINTER(C.c_float)]
        L.vrt_mfma_probe.argtypes = [vp, vp, vp, vp, ip]
        L.vrt_pathtrace_nrc.argtypes = [vp, C.POINTER(Scene), C.POINTER(Camera),
                                        C.POINTER(Params), C.POINTER(Frame), vp,
                                        C.POINTER(NrcRender), ip, ip]
        L.vrt_nrc_frame_counts.argtypes = [vp, C.POINTER(C.c_uint32)]
        L.vrt_relmse.argtypes = [vp, vp, vp, C.c_double, ip, C.c_float,
                                 C.POINTER(C.c_double)]
        L.vrt_sample_field.argtypes = [vp, C.POINTER(Volume), ip, vp, vp]
        L.vrt_nanovdb_create.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_void_p)]
        L.vrt_nanovdb_get_info.argtypes = [vp, C.c_void_p]
        L.vrt_nanovdb_destroy.argtypes = [vp]
        L.vrt_unstructured_create.argtypes = [vp, vp, vp, C.c_int64, vp, C.c_int64, vp, vp, vp,
                                              C.c_int64, C.POINTER(C.c_void_p)]
        L.vrt_unstructured_get_info.argtypes = [vp, C.c_void_p]
        L.vrt_unstructured_destroy.argtypes = [vp]
        L.vrt_intersect_surfaces.argtypes = [vp, C.POINTER(Scene), ip, vp, ip, vp, vp, vp]
        L.vrt_instance_transforms.argtypes = [vp, vp, vp, vp]
        L.vrt_tls_build.argtypes = [vp, vp, ip, C.POINTER(C.c_void_p)]
        L.vrt_tls_get_info.argtypes = [vp, C.c_void_p]
        L.vrt_tls_destroy.argtypes = [vp]
        L.vrt_bvh_build.argtypes = [vp, vp, C.c_int64, vp, C.c_int64, C.POINTER(C.c_void_p)]
        L.vrt_bvh_build_gpu.argtypes = [vp, vp, C.c_int64, vp, C.c_int64, C.POINTER(C.c_void_p)]
        L.vrt_bvh_get_info.argtypes = [vp, C.c_void_p]
        L.vrt_bvh_destroy.argtypes = [vp]
        L.vrt_hdri_create.argtypes = [vp, vp, ip, ip, C.POINTER(C.c_void_p)]
        L.vrt_hdri_destroy.argtypes = [vp]
        L.vrt_delta_track.argtypes = [vp, C.POINTER(Scene), ip, vp, vp, vp]
        L.vrt_raycast.argtypes = [vp, C.POINTER(Scene), C.POINTER(Camera), C.POINTER(Params),
                                  C.POINTER(Frame), ip]
        L.vrt_pathtrace.argtypes = [vp, C.POINTER(Scene), C.POINTER(Camera), C.POINTER(Params),
                                    C.POINTER(Frame), ip, ip]
        L.vrt_pathtrace_noalloc.argtypes = [vp, C.POINTER(Scene), C.POINTER(Camera),
                                            C.POINTER(Params), C.POINTER(Frame), ip, ip]
        _lib = L
    return _lib


def gen():
    """The loaded scene-generator library."""
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise VrtError(f"{GEN_PATH} is missing: run python anari-visionaray_amd/build.py")
        G = C.CDLL(GEN_PATH)
        G.vrtgen_blobs.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_uint64, C.c_void_p, C.c_int]
        G.vrtgen_fbm.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_uint64, C.c_void_p, C.c_int]
        G.vrtgen_splitmix_uniform.argtypes = [C.c_uint64, C.c_int, C.c_void_p]
        G.vrtgen_nanovdb_from_dense.argtypes = [C.POINTER(C.c_int), C.c_void_p, C.c_float,
                                                C.c_double, C.POINTER(C.c_double),
                                                C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        G.vrtgen_nanovdb_cloud.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                           C.c_uint64, C.c_double, C.c_int,
                                           C.POINTER(C.c_void_p), C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_double)]
        G.vrtgen_free.argtypes = [C.c_void_p]
        G.vrtgen_icosphere.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                       C.c_void_p, C.c_void_p]
        G.vrtgen_tet_lattice.argtypes = [C.c_int, C.c_double, C.c_int, C.c_uint64, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_int]
        _gen = G
    return _gen


def check(status):
    if status != 0:
        msg = lib().vrt_last_error().decode(errors="replace")
        raise VrtError(f"vrt status {status}: {msg}")
    return status


def _dims(d):
    return (C.c_int * 3)(*[int(x) for x in d])


def gen_blobs(dims, count=32, seed=0, threads=0):
    out = np.empty((dims[2], dims[1], dims[0]), np.float32)
    gen().vrtgen_blobs(_dims(dims), count, seed, out.ctypes.data, threads)
    return out


def gen_fbm(dims, octaves=6, base_freq=4.0, lacunarity=2.0, gain=0.5, seed=0, threads=0):
    out = np.empty((dims[2], dims[1], dims[0]), np.float32)
    gen().vrtgen_fbm(_dims(dims), octaves, base_freq, lacunarity, gain, seed, out.ctypes.data,
                     threads)
    return out


def aligned_bytes(n, align=64):
    """A uint8 array of n bytes whose data pointer is `align`-byte aligned
    (NanoVDB buffers need 32-byte alignment)."""
    raw = np.empty(int(n) + align, np.uint8)
    off = (-raw.ctypes.data) % align
    return raw[off:off + int(n)]


def _take_gen_buffer(ptr, nbytes):
    out = aligned_bytes(nbytes.value)
    C.memmove(out.ctypes.data, ptr.value, nbytes.value)
    gen().vrtgen_free(ptr)
    return out


def gen_nanovdb_from_dense(values, background=0.0, voxel=1.0, trans=(0.0, 0.0, 0.0)):
    """NanoVDB float grid of a dense (z, y, x) array: voxels != background are
    active; world = voxel * index + trans."""
    v = np.ascontiguousarray(values, np.float32)
    ptr, n = C.c_void_p(), C.c_uint64()
    dims = (v.shape[2], v.shape[1], v.shape[0])
    if gen().vrtgen_nanovdb_from_dense(_dims(dims), v.ctypes.data, background, voxel,
                                       (C.c_double * 3)(*trans), C.byref(ptr), C.byref(n)):
        raise VrtError("nanovdb writer failed")
    return _take_gen_buffer(ptr, n)


def gen_nanovdb_cloud(n=1024, octaves=6, base_freq=4.0, lacunarity=2.0, gain=0.5, seed=3,
                      active_fraction=0.05, threads=0):
    """Sparse fBm cloud as a NanoVDB float grid over index space [0, n)^3
    (identity map).  Returns (buffer, threshold)."""
    ptr, nb, tau = C.c_void_p(), C.c_uint64(), C.c_double()
    if gen().vrtgen_nanovdb_cloud(n, octaves, base_freq, lacunarity, gain, seed, active_fraction,
                                  threads, C.byref(ptr), C.byref(nb), C.byref(tau)):
        raise VrtError("nanovdb cloud generator failed")
    return _take_gen_buffer(ptr, nb), tau.value


def gen_icosphere(level=8, radius=0.5, amp=0.08, freq=3.0, seed=2):
    """Noise-displaced icosphere: (vertices (n, 3) float32, indices (m, 3) uint32),
    m = 20 * 4^level."""
    nt = 20 * 4 ** level
    nv = 10 * 4 ** level + 2
    v = np.empty((nv, 3), np.float32)
    i = np.empty((nt, 3), np.uint32)
    if gen().vrtgen_icosphere(level, radius, amp, freq, seed, v.ctypes.data, i.ctypes.data):
        raise VrtError("icosphere generator failed")
    return v, i


def splitmix_uniform(seed, n):
    """n uniform doubles in [0, 1) from SplitMix64(seed)."""
    out = np.empty(n, np.float64)
    gen().vrtgen_splitmix_uniform(seed, n, out.ctypes.data)
    return out


def gen_tet_lattice(n=70, jitter=0.15, blob_count=32, seed=5, threads=0):
    """Jittered lattice of n^3 cubes over [-0.5, 0.5]^3, 6 tetrahedra per cube,
    Gaussian-blob vertex values: (positions (V, 3) float32, values (V,) float32,
    index (T, 4) uint32), V = (n+1)^3, T = 6 n^3."""
    nv, nt = (n + 1) ** 3, 6 * n ** 3
    pos = np.empty((nv, 3), np.float32)
    val = np.empty(nv, np.float32)
    idx = np.empty((nt, 4), np.uint32)
    if gen().vrtgen_tet_lattice(n, jitter, blob_count, seed, pos.ctypes.data, val.ctypes.data,
                                idx.ctypes.data, threads):
        raise VrtError("tetrahedral lattice generator failed (degenerate cells)")
    return pos, val, idx


def tet_cells(num_tets):
    """cell.index / cell.type arrays of a tetrahedra-only index buffer."""
    return (np.arange(num_tets, dtype=np.uint32) * 4, np.full(num_tets, 10, np.uint8))


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().vrt_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes, (arr.nbytes, self.nbytes)
        check(lib().vrt_upload(self.ctx.handle, self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(lib().vrt_download(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self):
        check(lib().vrt_memset(self.ctx.handle, self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            check(lib().vrt_free(self.ctx.handle, self.ptr))
            self.ptr = None


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().vrt_context_create(device, C.byref(h)))
        self.handle = h.value
        self._bufs = []

    def buffer(self, nbytes):
        b = DeviceBuffer(self, nbytes)
        self._bufs.append(b)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        b = self.buffer(max(arr.nbytes, 4))
        b.upload(arr)
        return b

    def sync(self):
        check(lib().vrt_synchronize(self.handle))

    def copy(self, dst, src, nbytes):
        check(lib().vrt_device_copy(self.handle, dst.ptr, src.ptr, nbytes))

    def timer_start(self):
        check(lib().vrt_timer_start(self.handle))

    def timer_stop(self):
        ms = C.c_float()
        check(lib().vrt_timer_stop(self.handle, C.byref(ms)))
        return ms.value

    def close(self):
        for b in self._bufs:
            b.free()
        self._bufs = []
        if self.handle:
            check(lib().vrt_context_destroy(self.handle))
            self.handle = None


class NanoVDBInfo(C.Structure):
    _fields_ = [("bounds_lo", f3), ("bounds_hi", f3), ("background", C.c_float),
                ("top_origin", i3), ("top_dims", i3), ("blocks", C.c_uint64),
                ("leaves", C.c_uint64),
                ("constants", C.c_uint64), ("active_voxels", C.c_uint64),
                ("device_bytes", C.c_uint64)]


class DeviceNanoVDB:
    """A NanoVDB float grid converted to the GPU slot-table layout."""

    def __init__(self, ctx, grid_buffer):
        buf = np.ascontiguousarray(grid_buffer, np.uint8)
        h = C.c_void_p()
        check(lib().vrt_nanovdb_create(ctx.handle, buf.ctypes.data, buf.nbytes, C.byref(h)))
        self.handle = h.value
        self.info = NanoVDBInfo()
        check(lib().vrt_nanovdb_get_info(self.handle, C.byref(self.info)))

    def close(self):
        if self.handle:
            lib().vrt_nanovdb_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class UnstructuredInfo(C.Structure):
    _fields_ = [("bounds_lo", f3), ("bounds_hi", f3), ("cell_size", C.c_float),
                ("num_cells", C.c_uint64), ("skipped_cells", C.c_uint64), ("grid_dims", i3),
                ("list_entries", C.c_uint64), ("max_list", C.c_uint32),
                ("mean_list", C.c_double), ("expected_tests", C.c_double),
                ("build_ms", C.c_double), ("device_bytes", C.c_uint64)]


class DeviceUnstructured:
    """An unstructured (tetrahedral) field in the GPU point-location layout,
    from the field's ANARI arrays."""

    def __init__(self, ctx, positions, index, cell_index=None, cell_type=None,
                 vertex_data=None, cell_data=None):
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(index, np.uint32).ravel()
        if cell_index is None:
            cell_index, cell_type = tet_cells(len(idx) // 4)
        ci = np.ascontiguousarray(cell_index, np.uint32)
        ct = np.ascontiguousarray(cell_type, np.uint8)
        vd = None if vertex_data is None else np.ascontiguousarray(vertex_data, np.float32)
        cd = None if cell_data is None else np.ascontiguousarray(cell_data, np.float32)
        h = C.c_void_p()
        check(lib().vrt_unstructured_create(
            ctx.handle, pos.ctypes.data, None if vd is None else vd.ctypes.data, len(pos),
            idx.ctypes.data, len(idx), ci.ctypes.data, ct.ctypes.data,
            None if cd is None else cd.ctypes.data, len(ci), C.byref(h)))
        self.handle = h.value
        self.info = UnstructuredInfo()
        check(lib().vrt_unstructured_get_info(self.handle, C.byref(self.info)))

    def close(self):
        if self.handle:
            lib().vrt_unstructured_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceVolume:
    """A structuredRegular, nanovdb or unstructured field + transferFunction1D
    volume resident on the GPU (field data, TF table, value ranges and
    majorant grid)."""

    @classmethod
    def from_unstructured(cls, ctx, field, tf_rgba, value_range, unit_distance, grid_dims=None,
                          xfm=None):
        """`field`: a DeviceUnstructured."""
        self = cls.__new__(cls)
        self.field = field
        v = Volume()
        v.field_type = FIELD_UNSTRUCTURED
        v.unstructured = field.handle
        v.bounds_lo = field.info.bounds_lo
        v.bounds_hi = field.info.bounds_hi
        v.cell_size = field.info.cell_size
        self.dims = None
        self._finish(ctx, v, tf_rgba, value_range, unit_distance, grid_dims, xfm)
        return self

    @classmethod
    def from_nanovdb(cls, ctx, grid, tf_rgba, value_range, unit_distance, grid_dims=None,
                     xfm=None, filter="linear"):
        """`grid`: a NanoVDB buffer (uint8 array) or a DeviceNanoVDB."""
        self = cls.__new__(cls)
        self.grid = grid if isinstance(grid, DeviceNanoVDB) else DeviceNanoVDB(ctx, grid)
        v = Volume()
        v.field_type = FIELD_NANOVDB
        v.nanovdb = self.grid.handle
        v.filter = {"linear": 0, "nearest": 1}[filter]
        v.bounds_lo = self.grid.info.bounds_lo
        v.bounds_hi = self.grid.info.bounds_hi
        v.cell_size = 1.0
        self.dims = None
        self._finish(ctx, v, tf_rgba, value_range, unit_distance, grid_dims, xfm)
        return self

    def __init__(self, ctx, voxels, origin, spacing, tf_rgba, value_range, unit_distance,
                 grid_dims=None, xfm=None):
        voxels = np.ascontiguousarray(voxels, np.float32)
        nz, ny, nx = voxels.shape
        self.dims = (nx, ny, nz)
        lin = ctx.upload(voxels)
        nbytes = lib().vrt_structured_brick_bytes(_dims(self.dims))
        self.bricks = ctx.buffer(nbytes)
        check(lib().vrt_structured_bricks(ctx.handle, lin.ptr, _dims(self.dims), self.bricks.ptr))
        ctx.sync()
        lin.free()
        v = Volume()
        v.field_type = FIELD_STRUCTURED_REGULAR
        v.dims = i3(*self.dims)
        lo = np.array(origin, np.float32)
        hi = lo + (np.array(self.dims, np.float32) - np.float32(1)) * np.array(spacing, np.float32)
        v.bounds_lo = f3(*lo)
        v.bounds_hi = f3(*hi)
        v.cell_size = float(np.min(np.array(spacing, np.float32)))
        v.bricks = self.bricks.ptr
        self._finish(ctx, v, tf_rgba, value_range, unit_distance, grid_dims, xfm)

    def _finish(self, ctx, v, tf_rgba, value_range, unit_distance, grid_dims, xfm):
        v.object_id = NO_ID
        v.instance_id = NO_ID
        tf = np.ascontiguousarray(tf_rgba, np.float32).reshape(-1, 4)
        self.tf = ctx.upload(tf)
        v.tf_rgba = self.tf.ptr
        v.tf_size = tf.shape[0]
        v.value_lo, v.value_hi = float(value_range[0]), float(value_range[1])
        v.unit_distance = float(unit_distance)
        A, tinv = affine_inverse(xfm)
        v.affine_inv = (C.c_float * 9)(*A.ravel())
        v.trans_inv = f3(*tinv)
        self.grid_dims = None
        if grid_dims is not None:
            self.grid_dims = tuple(int(g) for g in grid_dims)
            v.grid_dims = i3(*self.grid_dims)
            nmc = int(np.prod(self.grid_dims))
            self.ranges = ctx.buffer(nmc * 8)
            self.majorants = ctx.buffer(nmc * 4)
            self.majorant_bytes = ctx.buffer(nmc)
            self.majorant_step = ctx.buffer(4)
            check(lib().vrt_build_value_ranges(ctx.handle, C.byref(v), self.ranges.ptr))
            v.majorants = self.majorants.ptr
            v.majorant_bytes = self.majorant_bytes.ptr
            v.majorant_step = self.majorant_step.ptr
            check(lib().vrt_compute_majorants(ctx.handle, C.byref(v), self.ranges.ptr,
                                              self.majorants.ptr, self.majorant_bytes.ptr,
                                              self.majorant_step.ptr))
            ctx.sync()
        self.desc = v

    def majorant_array(self):
        return self.majorants.download(np.float32, self.grid_dims[::-1])


def affine_inverse(xfm):
    """(affineInv, transInv) of an object->world row-major 3x4 transform."""
    if xfm is None:
        return np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    m = np.asarray(xfm, np.float64).reshape(3, 4)
    A = np.linalg.inv(m[:, :3]).astype(np.float32)
    return A, (-m[:, 3]).astype(np.float32)


def make_scene(volumes, lights=(), surfaces=(), tls=None):
    """`surfaces`: (DeviceMesh, matte color) pairs; `tls`: a DeviceTls of
    instanced meshes."""
    s = Scene()
    s.tls = tls.handle if tls is not None else None
    s.num_volumes = len(volumes)
    for i, v in enumerate(volumes):
        s.volumes[i] = v.desc
    s.num_lights = len(lights)
    for i, L in enumerate(lights):
        s.lights[i] = L
    s.num_surfaces = len(surfaces)
    for i, (mesh, color) in enumerate(surfaces):
        s.surfaces[i].bvh = mesh.handle
        s.surfaces[i].color = f3(*color)
        s.surfaces[i].object_id = NO_ID
        s.surfaces[i].instance_id = NO_ID
    return s


class DeviceMesh:
    """A triangle mesh's 4-wide BVH resident on the GPU, built on the host
    by binned SAH (builder "sah") or on the GPU as a linear BVH ("lbvh")."""

    def __init__(self, ctx, vertices, indices=None, builder="sah"):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        ip, nt, idx = None, len(v) // 3, None
        if indices is not None:
            idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
            ip, nt = idx.ctypes.data, len(idx)
        h = C.c_void_p()
        if builder == "lbvh":
            dv = ctx.upload(v)
            di = ctx.upload(idx) if idx is not None else None
            ctx.sync()
            try:
                check(lib().vrt_bvh_build_gpu(ctx.handle, dv.ptr, len(v),
                                              di.ptr if di is not None else None, nt,
                                              C.byref(h)))
            finally:
                dv.free()
                if di is not None:
                    di.free()
        elif builder == "sah":
            check(lib().vrt_bvh_build(ctx.handle, v.ctypes.data, len(v), ip, nt, C.byref(h)))
        else:
            raise ValueError(f"unknown BVH builder {builder!r}")
        self.handle = h.value
        self.info = BvhInfo()
        check(lib().vrt_bvh_get_info(self.handle, C.byref(self.info)))

    def close(self):
        if self.handle:
            lib().vrt_bvh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def instance_transforms(transform):
    """(affine_inv, trans_inv, normal_xfm) of a row-major 3x4 [L | t] as the
    library derives them (vrt_instance_transforms)."""
    x = np.ascontiguousarray(transform, np.float32).reshape(12)
    a, t, n = np.empty(9, np.float32), np.empty(3, np.float32), np.empty(9, np.float32)
    check(lib().vrt_instance_transforms(x.ctypes.data, a.ctypes.data, t.ctypes.data,
                                        n.ctypes.data))
    return a.reshape(3, 3), t, n.reshape(3, 3)


class DeviceTls:
    """Instance-level BVH over meshes placed by transforms: `instances` are
    (DeviceMesh, row-major 3x4 transform, matte color[, object id, instance
    id]) tuples; the meshes must outlive it."""

    def __init__(self, ctx, instances):
        arr = (Instance * len(instances))()
        for i, inst in enumerate(instances):
            mesh, xf, color = inst[:3]
            arr[i].bvh = mesh.handle
            arr[i].transform = (C.c_float * 12)(*np.asarray(xf, np.float32).reshape(12))
            arr[i].color = f3(*color)
            arr[i].object_id = inst[3] if len(inst) > 3 else NO_ID
            arr[i].instance_id = inst[4] if len(inst) > 4 else NO_ID
        h = C.c_void_p()
        check(lib().vrt_tls_build(ctx.handle, arr, len(instances), C.byref(h)))
        self.handle = h.value
        self.info = TlsInfo()
        check(lib().vrt_tls_get_info(self.handle, C.byref(self.info)))

    def close(self):
        if self.handle:
            lib().vrt_tls_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def quad_light(position, edge1, edge2, intensity=1.0, color=(1, 1, 1), visible=True):
    L = Light()
    L.type = LIGHT_QUAD
    L.visible = int(visible)
    L.position = f3(*position)
    L.edge1 = f3(*edge1)
    L.edge2 = f3(*edge2)
    L.color = f3(*color)
    L.intensity = float(intensity)
    return L


class DeviceHdri:
    """An HDRI environment map (h, w, 3) float32 on the GPU with its
    sampling CDFs (vrt_hdri_create)."""

    def __init__(self, ctx, radiance):
        rgb = np.ascontiguousarray(radiance, np.float32)
        h, w = rgb.shape[:2]
        out = C.c_void_p()
        check(lib().vrt_hdri_create(ctx.handle, rgb.ctypes.data, w, h, C.byref(out)))
        self.handle = out.value
        self.width, self.height = w, h

    def close(self):
        if self.handle:
            lib().vrt_hdri_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hdri_light(hdri, up=(1, 0, 0), direction=(0, 0, 1), scale=1.0, visible=True):
    """ANARI "hdri" light (HDRI.cpp:102-108 defaults)."""
    L = Light()
    L.type = LIGHT_HDRI
    L.visible = int(visible)
    L.hdri = hdri.handle
    L.up = f3(*up)
    L.direction = f3(*direction)
    L.scale = float(scale)
    L.color = f3(1, 1, 1)
    return L


def directional_light(direction, irradiance=1.0, color=(1, 1, 1), visible=True):
    L = Light()
    L.type = LIGHT_DIRECTIONAL
    L.visible = int(visible)
    L.direction = f3(*direction)
    L.color = f3(*color)
    L.intensity = float(irradiance)
    return L


def camera(position, direction, up, fovy, aspect, region=(0, 0, 1, 1), shutter=(0.5, 0.5),
           aperture_radius=0.0, focus_distance=1.0):
    c = Camera()
    c.position = f3(*position)
    c.direction = f3(*direction)
    c.up = f3(*up)
    c.fovy = float(fovy)
    c.aspect = float(aspect)
    c.image_region = f4(*region)
    c.shutter = (C.c_float * 2)(*shutter)
    c.aperture_radius = float(aperture_radius)
    c.focus_distance = float(focus_distance)
    return c


def params(background=(0, 0, 0, 1), ambient_color=(1, 1, 1), ambient_radiance=0.2,
           occlusion_distance=1e20, ambient_samples=1, max_bounce=7, pixel_samples=1,
           volume_sampling_rate=0.5, rng_offset=0, clip_planes=(), phase_g=0.0,
           gradient_shading=False, taa=False, taa_alpha=0.3, prev_camera=None,
           ratio_tracking=False, pass_schedule=0):
    """Renderer parameters with the ANARI defaults of the reference
    (Renderer.cpp:42-53, Pathtrace.cpp:19-25, Raycast.cpp:19).  `clip_planes`:
    up to 8 (nx, ny, nz, d); rays keep n.x <= d.  `phase_g`: Henyey-Greenstein
    anisotropy of the volume phase (0: the reference's isotropic phase).
    `ratio_tracking`: shadow and AO transmittance by ratio tracking (the
    reference: binary delta tracking)."""
    p = Params()
    p.ratio_tracking = int(bool(ratio_tracking))
    p.pass_schedule = int(pass_schedule)
    p.rng_offset = rng_offset
    p.phase_g = phase_g
    p.gradient_shading = int(bool(gradient_shading))
    p.taa = int(bool(taa))
    p.taa_alpha = taa_alpha
    if prev_camera is not None:  # a Camera from camera()
        p.has_prev_camera = 1
        p.prev_camera = prev_camera
    if len(clip_planes) > 8:
        raise ValueError("at most 8 clip planes")
    p.num_clip_planes = len(clip_planes)
    for i, c in enumerate(clip_planes):
        p.clip_planes[i] = (C.c_float * 4)(*c)
    p.background = f4(*background)
    p.ambient_color = f3(*ambient_color)
    p.ambient_radiance = ambient_radiance
    p.occlusion_distance = occlusion_distance
    p.ambient_samples = ambient_samples
    p.max_bounce = max_bounce
    p.pixel_samples = pixel_samples
    p.volume_sampling_rate = volume_sampling_rate
    return p


class DeviceFrame:
    def __init__(self, ctx, width, height, color_type=TYPE_FLOAT32_VEC4, depth=False,
                 accum64=False, counters=True, gbuffer=False, taa=False):
        """gbuffer: also the normal, albedo and primitive / object / instance
        id channels (they need the depth channel, which it turns on)."""
        self.ctx = ctx
        self.width, self.height = width, height
        n = width * height
        bpp = 16 if color_type == TYPE_FLOAT32_VEC4 else 4
        self.color = ctx.buffer(n * bpp)
        self.accum = ctx.buffer(n * 16)
        self.depth = ctx.buffer(n * 4) if depth else None
        self.accum64 = ctx.buffer(n * 32) if accum64 else None
        self.counters = ctx.buffer(64) if counters else None
        if gbuffer and not depth:
            depth = True
            self.depth = ctx.buffer(n * 4)
        self.gbuf = {k: ctx.buffer(n * b) for k, b in
                     (("normal", 12), ("albedo", 12), ("primitive_id", 4), ("object_id", 4),
                      ("instance_id", 4))} if gbuffer else {}
        f = Frame()
        f.width, f.height, f.color_type = width, height, color_type
        f.color = self.color.ptr
        f.accum = self.accum.ptr
        f.depth = self.depth.ptr if depth else None
        f.accum64 = self.accum64.ptr if accum64 else None
        f.counters = self.counters.ptr if counters else None
        for k, b in self.gbuf.items():
            setattr(f, k, b.ptr)
        # temporal anti-aliasing: motion vectors, current / history colour and
        # albedo, zeroed (the reference resizes them with zeros)
        self.taa = {k: ctx.buffer(n * b) for k, b in
                    (("motion_vec", 16), ("taa_curr", 16), ("taa_prev", 16),
                     ("taa_curr_albedo", 12), ("taa_prev_albedo", 12))} if taa else {}
        for k, b in self.taa.items():
            b.zero()
            setattr(f, k, b.ptr)
        self.desc = f
        self.reset()

    def reset(self):
        check(lib().vrt_frame_reset(self.ctx.handle, C.byref(self.desc)))
        if self.counters:
            self.counters.zero()

    def color_array(self):
        if self.desc.color_type == TYPE_FLOAT32_VEC4:
            return self.color.download(np.float32, (self.height, self.width, 4))
        return self.color.download(np.uint32, (self.height, self.width))

    def accum_array(self):
        return self.accum.download(np.float32, (self.height, self.width, 4))

    def accum64_array(self):
        return self.accum64.download(np.float64, (self.height, self.width, 4))

    def depth_array(self):
        return self.depth.download(np.float32, (self.height, self.width))

    def gbuffer_arrays(self):
        """{normal, albedo: (h, w, 3) float32; primitive_id, object_id,
        instance_id: (h, w) uint32}"""
        out = {}
        for k, b in self.gbuf.items():
            if k in ("normal", "albedo"):
                out[k] = b.download(np.float32, (self.height, self.width, 3))
            else:
                out[k] = b.download(np.uint32, (self.height, self.width))
        return out

    def counter_values(self):
        return self.counters.download(np.uint64, (8,))


def sample_field(ctx, volume, points):
    """Trilinear values of a DeviceVolume at voxel-space points (n, 3)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = pts.shape[0]
    dp = ctx.upload(pts)
    dv = ctx.buffer(max(4, n * 4))
    check(lib().vrt_sample_field(ctx.handle, C.byref(volume.desc), n, dp.ptr, dv.ptr))
    out = dv.download(np.float32, (n,))
    dp.free()
    dv.free()
    return out


def intersect_surfaces(ctx, scene, rays, any_hit=False, which=False):
    """Surface hits of rays (n, 8): {ori, dir, tmin, tmax} -> (t, prim), and
    with which=True also the surface hit (surface i, or num_surfaces + j for
    instance j of the scene's TLS; -1 on a miss)."""
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = len(r)
    dr = ctx.upload(r)
    dt, dp, ds = ctx.buffer(max(4, 4 * n)), ctx.buffer(max(4, 4 * n)), ctx.buffer(max(4, 4 * n))
    check(lib().vrt_intersect_surfaces(ctx.handle, C.byref(scene), n, dr.ptr, int(any_hit),
                                       dt.ptr, dp.ptr, ds.ptr if which else None))
    t = dt.download(np.float32, (n,))
    p = dp.download(np.int32, (n,))
    sf = ds.download(np.int32, (n,)) if which else None
    for b in (dr, dt, dp, ds):
        b.free()
    return (t, p, sf) if which else (t, p)


def delta_track(ctx, scene, rays, rng_states):
    """Collision distances of world rays (n, 8) through scene volume 0."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    states = np.ascontiguousarray(rng_states, np.uint32)
    dr = ctx.upload(rays)
    ds = ctx.upload(states)
    dt = ctx.buffer(max(4, n * 4))
    check(lib().vrt_delta_track(ctx.handle, C.byref(scene), n, dr.ptr, ds.ptr, dt.ptr))
    t = dt.download(np.float32, (n,))
    s = ds.download(np.uint32, (n,))
    for b in (dr, ds, dt):
        b.free()
    return t, s


def raycast(ctx, scene, cam, prm, frame, frame_id=0):
    check(lib().vrt_raycast(ctx.handle, C.byref(scene), C.byref(cam), C.byref(prm),
                            C.byref(frame.desc), frame_id))


def pathtrace(ctx, scene, cam, prm, frame, accum_id, frame_id=0):
    check(lib().vrt_pathtrace(ctx.handle, C.byref(scene), C.byref(cam), C.byref(prm),
                              C.byref(frame.desc), accum_id, frame_id))


def pathtrace_noalloc(ctx, scene, cam, prm, frame, accum_id, frame_id=0):
    """vrt_pathtrace_noalloc: the same frame with no device allocation."""
    check(lib().vrt_pathtrace_noalloc(ctx.handle, C.byref(scene), C.byref(cam), C.byref(prm),
                                      C.byref(frame.desc), accum_id, frame_id))


class Nrc:
    """The neural radiance cache: fp32 master weights + Adam state on the GPU."""

    def __init__(self, ctx, precision=NRC_FP16, learning_rate=1e-2, beta1=0.9, beta2=0.99,
                 epsilon=1e-8, max_batch=65536, seed=0):
        self.ctx = ctx
        cfg = NrcConfig(precision, learning_rate, beta1, beta2, epsilon, max_batch, seed)
        h = C.c_void_p()
        check(lib().vrt_nrc_create(ctx.handle, C.byref(cfg), C.byref(h)))
        self.handle = h.value
        self.nparams = lib().vrt_nrc_param_count()

    def params(self):
        out = np.empty(self.nparams, np.float32)
        check(lib().vrt_nrc_get_params(self.handle, out.ctypes.data))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        check(lib().vrt_nrc_set_params(self.handle, p.ctypes.data))

    def infer(self, raw):
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 6)
        n = raw.shape[0]
        dr = self.ctx.upload(raw)
        do = self.ctx.buffer(max(4, n * 12))
        check(lib().vrt_nrc_infer(self.handle, n, dr.ptr, do.ptr))
        out = do.download(np.float32, (n, 3))
        dr.free()
        do.free()
        return out

    def gradients(self, raw, target):
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 6)
        target = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        n = raw.shape[0]
        dr, dt = self.ctx.upload(raw), self.ctx.upload(target)
        dg = self.ctx.buffer(self.nparams * 4)
        loss = C.c_float()
        check(lib().vrt_nrc_gradients(self.handle, n, dr.ptr, dt.ptr, dg.ptr, C.byref(loss)))
        g = dg.download(np.float32, (self.nparams,))
        for b in (dr, dt, dg):
            b.free()
        return g, loss.value

    def train_step(self, raw_buf, target_buf, n):
        """One Adam step on n device-resident samples; returns the loss sum."""
        loss = C.c_float()
        check(lib().vrt_nrc_train_step(self.handle, n, raw_buf.ptr, target_buf.ptr,
                                       C.byref(loss)))
        return loss.value

    def frame_counts(self):
        """[queries, training claims, vertex-2 paths, previous vertex-2 paths]"""
        out = (C.c_uint32 * 4)()
        check(lib().vrt_nrc_frame_counts(self.handle, out))
        return list(out)

    def close(self):
        if self.handle:
            check(lib().vrt_nrc_destroy(self.handle))
            self.handle = None


def nrc_render_options(bounds_lo, bounds_hi, train=True, train_samples=65536, train_steps=4,
                       query_vertex=2, train_extra_vertices=0):
    o = NrcRender()
    o.query_vertex = int(query_vertex)
    o.train_extra_vertices = int(train_extra_vertices)
    o.train = int(train)
    o.train_samples = int(train_samples)
    o.train_steps = int(train_steps)
    o.bounds_lo = f3(*bounds_lo)
    o.bounds_hi = f3(*bounds_hi)
    return o


def relmse(ctx, frame, ref_frame, ref_frames, eps=1e-2):
    """relMSE of frame's running mean against ref_frame's FP64 sums / ref_frames."""
    out = C.c_double()
    check(lib().vrt_relmse(ctx.handle, frame.accum.ptr, ref_frame.accum64.ptr,
                           1.0 / ref_frames, frame.width * frame.height, eps, C.byref(out)))
    return out.value


def pathtrace_nrc(ctx, scene, cam, prm, frame, nrc, opts, accum_id, frame_id=0):
    check(lib().vrt_pathtrace_nrc(ctx.handle, C.byref(scene), C.byref(cam), C.byref(prm),
                                  C.byref(frame.desc), nrc.handle, C.byref(opts), accum_id,
                                  frame_id))


def mfma_probe(ctx, a, b, precision=NRC_FP16):
    """C (16x16) = A (16x32) B (32x16) through one 16x16x32 MFMA."""
    da = ctx.upload(np.ascontiguousarray(a, np.float32))
    db = ctx.upload(np.ascontiguousarray(b, np.float32))
    dc = ctx.buffer(16 * 16 * 4)
    check(lib().vrt_mfma_probe(ctx.handle, da.ptr, db.ptr, dc.ptr, precision))
    c = dc.download(np.float32, (16, 16))
    for x in (da, db, dc):
        x.free()
    return c
