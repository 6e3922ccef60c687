import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "trixicuda.jl_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950)")


def load_trixi_hip():
    """Import the host package that lives in ``trixicuda.jl_amd/`` as ``trixi_hip``."""
    if "trixi_hip" in sys.modules:
        return sys.modules["trixi_hip"]
    spec = importlib.util.spec_from_file_location(
        "trixi_hip", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["trixi_hip"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def th():
    return load_trixi_hip()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O
