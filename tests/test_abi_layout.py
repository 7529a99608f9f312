"""The ctypes mirrors in vrt.py against include/vrt.h: a small C program built
with the host compiler prints sizeof and every field offset of each public
struct, and each must equal the mirror's (a field added on one side only
shifts what the library reads)."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = [("Volume", "vrt_volume"), ("Light", "vrt_light"), ("Surface", "vrt_surface"),
         ("Scene", "vrt_scene"), ("BvhInfo", "vrt_bvh_info"), ("Camera", "vrt_camera"),
         ("Params", "vrt_params"), ("Frame", "vrt_frame"), ("NrcConfig", "vrt_nrc_config"),
         ("NrcRender", "vrt_nrc_render"), ("NanoVDBInfo", "vrt_nanovdb_info"),
         ("UnstructuredInfo", "vrt_unstructured_info"), ("Instance", "vrt_instance"),
         ("TlsInfo", "vrt_tls_info")]


@pytest.fixture(scope="module")
def c_layout(pkg, tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "vrt.h"', "int main(void) {"]
    for py, cn in PAIRS:
        cls = getattr(pkg.vrt, py)
        lines.append(f'  printf("{py} size %zu\\n", sizeof({cn}));')
        for name, *_ in cls._fields_:
            lines.append(f'  printf("{py} {name} %zu\\n", offsetof({cn}, {name}));')
    lines += ["  return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.splitlines():
        py, key, val = line.split()
        layout[(py, key)] = int(val)
    return layout


@pytest.mark.parametrize("py,cn", PAIRS)
def test_struct_matches_header(pkg, c_layout, py, cn):
    cls = getattr(pkg.vrt, py)
    assert C.sizeof(cls) == c_layout[(py, "size")], f"{py} vs {cn}: size"
    for name, *_ in cls._fields_:
        assert getattr(cls, name).offset == c_layout[(py, name)], f"{py}.{name} vs {cn}.{name}"
