import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "ls-stag-3d_parallel_amd")
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def load_lsx():
    """Import the product's Python mirror (ls-stag-3d_parallel_amd/lsx.py)."""
    name = "lsx_pkg"
    if name in sys.modules:
        return sys.modules[name].lsx
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod.lsx


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def lsx():
    return load_lsx()
