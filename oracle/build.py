"""ORACLE build recipe -- test infrastructure.

Compiles the CPU restatement (vrt_oracle.cpp) into oracle/_build/.
-ffp-contract=off keeps every float operation individually rounded, and
x86-64-v3 (AVX2/FMA) runs on any current EPYC host.

When the reference tree is present, also compiles nvdb_ref.cpp against the
reference's vendored NanoVDB headers (/root/reference/external/nanovdb)
into oracle/_ref/libnvdb_ref.so, the golden source for NanoVDB grids and
lookups (baseline x86-64 and -ffp-contract=off: no contraction, like the
restatement).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
REF_OUT = os.path.join(HERE, "_ref")
REFERENCE_EXTERNAL = "/root/reference/external"


def build():
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libvrt_oracle.so")
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-std=c++17",
           "-fPIC", "-shared", "-pthread", os.path.join(HERE, "vrt_oracle.cpp"),
           os.path.join(HERE, "nrc_oracle.cpp"), os.path.join(HERE, "nvdb_oracle.cpp"), "-o", lib]
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    build_ref()
    return lib


def build_ref():
    if not os.path.isdir(os.path.join(REFERENCE_EXTERNAL, "nanovdb")):
        print("reference NanoVDB headers absent: oracle/_ref not rebuilt", flush=True)
        return None
    os.makedirs(REF_OUT, exist_ok=True)
    lib = os.path.join(REF_OUT, "libnvdb_ref.so")
    cmd = ["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-fPIC", "-shared",
           "-I" + REFERENCE_EXTERNAL, os.path.join(HERE, "nvdb_ref.cpp"), "-o", lib]
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return lib


if __name__ == "__main__":
    build()
