import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "_build", "libvrt_oracle.so")
    if not os.path.exists(lib):
        import importlib.util
        spec = importlib.util.spec_from_file_location(
            "oracle_build", os.path.join(ROOT, "oracle", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()


@pytest.fixture(scope="session")
def pkg():
    """The product package (anari-visionaray_amd/)."""
    sys.path.insert(0, ROOT)
    import importlib.util
    d = os.path.join(ROOT, "anari-visionaray_amd")
    name = "anari_visionaray_amd"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(
            name, os.path.join(d, "__init__.py"), submodule_search_locations=[d])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    import anari_visionaray_amd.vrt  # noqa: F401
    import anari_visionaray_amd.workloads  # noqa: F401
    return sys.modules[name]


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    import oracle_lib
    return oracle_lib


@pytest.fixture(scope="session")
def gpu_ctx(pkg):
    vrt = pkg.vrt
    ctx = vrt.Context(0)
    yield ctx
    ctx.close()
