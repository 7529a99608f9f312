"""ORACLE build recipe -- test infrastructure.

Compiles the CPU restatement (vrt_oracle.cpp) into oracle/_build/.
-ffp-contract=off keeps every float operation individually rounded, and
x86-64-v3 (AVX2/FMA) runs on any current EPYC host.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")


def build():
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libvrt_oracle.so")
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-std=c++17",
           "-fPIC", "-shared", "-pthread", os.path.join(HERE, "vrt_oracle.cpp"),
           os.path.join(HERE, "nrc_oracle.cpp"), "-o", lib]
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return lib


if __name__ == "__main__":
    build()
