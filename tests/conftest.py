import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mass-prime-siever_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_mps():
    """The product's Python view of libmps.so (the package directory name is not importable)."""
    if "mps" in sys.modules:
        return sys.modules["mps"]
    spec = importlib.util.spec_from_file_location("mps", os.path.join(PKG, "mps.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mps"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def mps():
    return load_mps()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu(mps):
    s = mps.GpuSiever()
    assert s.init(0), s.banner
    print("\n" + s.banner, flush=True)
    yield s
    s.close()
