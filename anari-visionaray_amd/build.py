"""Build the native libraries in-tree.

libvrt.so          gfx950 HIP kernels + the C ABI of include/vrt.h
libanari_vrt.so    the ANARI C API front end of include/anari_vrt.h, over libvrt.so
libvrt_scenegen.so host-only seeded scene generators (inputs for tests/bench)

Run as ``python anari-visionaray_amd/build.py``; every call recompiles.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: keep the compiler from forming v_pk_*_f32 instructions.
# -ffp-contract=off: every float operation of the rendering path is rounded
# on its own, as in the CPU restatement, so raycast results are
# bit-comparable; fused multiply-adds are written out explicitly (fmaf)
# where they are wanted.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC,
]

HIP_SOURCES = ["render_kernels.hip", "nrc_kernels.hip", "vrt_api.cpp", "nrc_api.cpp",
               "nanovdb_field.cpp", "bvh_build.cpp", "bvh_gpu.hip", "unstructured_field.cpp",
               "hdri_light.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_vrt():
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, "-x", "hip", "-c", path, "-o", obj] + HIP_FLAGS)
        objs.append(obj)
    lib = os.path.join(OUT, "libvrt.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", lib] + objs)
    return lib


def build_anari():
    os.makedirs(OUT, exist_ok=True)
    obj = os.path.join(OUT, "anari_device.o")
    _run([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, "anari_device.cpp"), "-o", obj] + HIP_FLAGS)
    lib = os.path.join(OUT, "libanari_vrt.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", lib, obj, "-L", OUT, "-lvrt",
          "-Wl,-rpath,$ORIGIN"])
    return lib


def build_scenegen():
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libvrt_scenegen.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread",
          os.path.join(CSRC, "scenegen.cpp"), "-o", lib])
    return lib


def build_all():
    build_scenegen()
    build_vrt()
    build_anari()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
