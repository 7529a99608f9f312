"""ctypes binding of the ANARI C API subset (include/anari_vrt.h) for tests:
scenes are built the way an ANARI application builds them."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "anari-visionaray_amd", "_build", "libanari_vrt.so")

STRING, BOOL, DATA_TYPE = 101, 103, 100
ARRAY1D, ARRAY2D, ARRAY3D = 504, 505, 506
CAMERA, FRAME, GEOMETRY, GROUP, INSTANCE, LIGHT = 507, 508, 509, 510, 511, 512
MATERIAL, RENDERER, SURFACE, SPATIAL_FIELD, VOLUME, WORLD = 513, 514, 515, 517, 518, 519
UINT8, INT32, UINT32, UINT32_VEC2, UINT32_VEC3 = 1004, 1016, 1020, 1021, 1022
UFIXED8_VEC4, FLOAT32, FLOAT32_VEC2, FLOAT32_VEC3, FLOAT32_VEC4 = 1039, 1068, 1069, 1070, 1071
FLOAT32_BOX1, FLOAT32_BOX2, FLOAT32_MAT4 = 2008, 2009, 2015
VOID_POINTER, FRAME_COMPLETION_CALLBACK = 102, 203
NO_WAIT, WAIT = 0, 1
FRAME_CALLBACK = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)

_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                  C.c_char_p)
_NP = {UINT8: (np.uint8, 1), INT32: (np.int32, 1), UINT32: (np.uint32, 1),
       UINT32_VEC2: (np.uint32, 2), UINT32_VEC3: (np.uint32, 3), FLOAT32: (np.float32, 1),
       FLOAT32_VEC2: (np.float32, 2), FLOAT32_VEC3: (np.float32, 3),
       FLOAT32_VEC4: (np.float32, 4), FLOAT32_BOX1: (np.float32, 2),
       FLOAT32_BOX2: (np.float32, 4), FLOAT32_MAT4: (np.float32, 16), BOOL: (np.int32, 1),
       DATA_TYPE: (np.int32, 1)}


class Anari:
    def __init__(self):
        L = C.CDLL(LIB_PATH)
        vp, u64, ip = C.c_void_p, C.c_uint64, C.c_int
        L.anariNewVisionarayDevice.restype = vp
        L.anariNewVisionarayDevice.argtypes = [_CB, vp]
        for fn, n in (("anariNewArray1D", 1), ("anariNewArray2D", 2), ("anariNewArray3D", 3)):
            getattr(L, fn).restype = vp
            getattr(L, fn).argtypes = [vp, vp, vp, vp, ip] + [u64] * n
        for fn in ("anariNewGeometry", "anariNewMaterial", "anariNewSpatialField",
                   "anariNewVolume", "anariNewLight", "anariNewInstance", "anariNewCamera",
                   "anariNewRenderer"):
            getattr(L, fn).restype = vp
            getattr(L, fn).argtypes = [vp, C.c_char_p]
        for fn in ("anariNewSurface", "anariNewGroup", "anariNewWorld", "anariNewFrame"):
            getattr(L, fn).restype = vp
            getattr(L, fn).argtypes = [vp]
        L.anariSetParameter.argtypes = [vp, vp, C.c_char_p, ip, vp]
        L.anariUnsetParameter.argtypes = [vp, vp, C.c_char_p]
        L.anariCommitParameters.argtypes = [vp, vp]
        L.anariRelease.argtypes = [vp, vp]
        L.anariRetain.argtypes = [vp, vp]
        L.anariGetProperty.argtypes = [vp, vp, C.c_char_p, ip, vp, u64, ip]
        L.anariRenderFrame.argtypes = [vp, vp]
        L.anariFrameReady.argtypes = [vp, vp, ip]
        L.anariMapFrame.restype = vp
        L.anariMapFrame.argtypes = [vp, vp, C.c_char_p, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_int)]
        L.anariUnmapFrame.argtypes = [vp, vp, C.c_char_p]
        L.anariMapArray.restype = vp
        L.anariMapArray.argtypes = [vp, vp]
        L.anariUnmapArray.argtypes = [vp, vp]
        self.L = L
        self.messages = []
        self._cb = _CB(lambda ud, dev, src, st, sev, code, msg:
                       self.messages.append((sev, code, msg.decode())))
        self.dev = L.anariNewVisionarayDevice(self._cb, None)
        if not self.dev:
            raise RuntimeError(f"no ANARI device: {self.messages}")
        self._keep = []

    # objects ---------------------------------------------------------------
    def new(self, kind, subtype=None):
        fn = {GEOMETRY: "anariNewGeometry", MATERIAL: "anariNewMaterial",
              SPATIAL_FIELD: "anariNewSpatialField", VOLUME: "anariNewVolume",
              LIGHT: "anariNewLight", INSTANCE: "anariNewInstance", CAMERA: "anariNewCamera",
              RENDERER: "anariNewRenderer", SURFACE: "anariNewSurface", GROUP: "anariNewGroup",
              WORLD: "anariNewWorld", FRAME: "anariNewFrame"}[kind]
        f = getattr(self.L, fn)
        return f(self.dev, subtype.encode()) if subtype is not None else f(self.dev)

    def array1d(self, data, elem):
        a = np.ascontiguousarray(data)
        self._keep.append(a)
        n = a.size // _NP[elem][1] if elem in _NP else len(a)
        return self.L.anariNewArray1D(self.dev, a.ctypes.data, None, None, elem, n)

    def array2d(self, image, elem):
        a = np.ascontiguousarray(image)
        self._keep.append(a)
        h, w = a.shape[:2]
        return self.L.anariNewArray2D(self.dev, a.ctypes.data, None, None, elem, w, h)

    def array3d(self, voxels, elem=FLOAT32):
        a = np.ascontiguousarray(voxels)
        self._keep.append(a)
        nz, ny, nx = a.shape
        return self.L.anariNewArray3D(self.dev, a.ctypes.data, None, None, elem, nx, ny, nz)

    def objects(self, handles, kind):
        h = np.array([int(x) for x in handles], np.uint64)
        self._keep.append(h)
        return self.L.anariNewArray1D(self.dev, h.ctypes.data, None, None, kind, len(h))

    def set(self, obj, name, kind, value):
        if kind == STRING:
            buf = C.create_string_buffer(value.encode())
            self.L.anariSetParameter(self.dev, obj, name.encode(), kind, buf)
        elif kind in _NP:
            v = np.ascontiguousarray(np.asarray(value, _NP[kind][0]).ravel())
            self.L.anariSetParameter(self.dev, obj, name.encode(), kind, v.ctypes.data)
        else:  # an object handle
            h = C.c_void_p(value)
            self.L.anariSetParameter(self.dev, obj, name.encode(), kind, C.byref(h))

    def commit(self, *objs):
        for o in objs:
            self.L.anariCommitParameters(self.dev, o)

    def release(self, *objs):
        for o in objs:
            self.L.anariRelease(self.dev, o)

    # frames ----------------------------------------------------------------
    def render(self, frame, wait=True):
        self.L.anariRenderFrame(self.dev, frame)
        if wait:
            self.L.anariFrameReady(self.dev, frame, WAIT)

    def map(self, frame, channel="channel.color"):
        w, h, t = C.c_uint32(), C.c_uint32(), C.c_int()
        p = self.L.anariMapFrame(self.dev, frame, channel.encode(), C.byref(w), C.byref(h),
                                 C.byref(t))
        if not p:
            return None, t.value
        if t.value == FLOAT32_VEC4:
            n, dt, comp = w.value * h.value * 4, np.float32, 4
        elif t.value == FLOAT32_VEC3:
            n, dt, comp = w.value * h.value * 3, np.float32, 3
        elif t.value == FLOAT32:
            n, dt, comp = w.value * h.value, np.float32, 1
        elif t.value == UINT32:
            n, dt, comp = w.value * h.value, np.uint32, 1
        else:
            n, dt, comp = w.value * h.value * 4, np.uint8, 4
        buf = (C.c_uint8 * (n * np.dtype(dt).itemsize)).from_address(p)
        out = np.frombuffer(buf, dt).copy().reshape(h.value, w.value, comp)
        self.L.anariUnmapFrame(self.dev, frame, channel.encode())
        return (out[..., 0] if comp == 1 else out), t.value

    def prop_f(self, obj, name):
        v = C.c_float()
        ok = self.L.anariGetProperty(self.dev, obj, name.encode(), FLOAT32, C.byref(v), 4, WAIT)
        return v.value if ok else None

    def prop_i(self, obj, name):
        v = C.c_int32()
        ok = self.L.anariGetProperty(self.dev, obj, name.encode(), INT32, C.byref(v), 4, WAIT)
        return v.value if ok else None

    def close(self):
        if self.dev:
            self.L.anariRelease(self.dev, self.dev)
            self.dev = None
