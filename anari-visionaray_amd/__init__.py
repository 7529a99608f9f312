"""MI355X (gfx950) implementation of anari-visionaray's per-pixel rendering
path: HIP kernels behind the C ABI of include/vrt.h (libvrt.so), the ANARI
device front end (libanari_library_visionaray_amd.so), and Python bindings
used by the tests and the benchmark.

The directory name contains a hyphen; import it through ``load()`` below or
register it under the module name ``anari_visionaray_amd``.
"""
import importlib.util
import os
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))


def load():
    """Import this directory as the package ``anari_visionaray_amd``."""
    name = "anari_visionaray_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod
