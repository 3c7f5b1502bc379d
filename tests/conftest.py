import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "3d-ray_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on HIP)")


def _ensure_oracle():
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)


def _ensure_lib():
    so = os.path.join(ROOT, "3d-ray_amd", "lib3dray_amd.so")
    if not os.path.exists(so):
        import importlib.util
        spec = importlib.util.spec_from_file_location("rt3d_build", os.path.join(ROOT, "3d-ray_amd", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def native():
    _ensure_lib()
    from raytracer import native as N
    N.load()
    return N


@pytest.fixture(scope="session")
def gpu_ctx(native):
    ctx = native.Context(0)
    yield ctx
    ctx.close()
