"""Shared test setup: load the package (its directory name is not an
identifier), register the `gpu` marker, and provide seeded workloads."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "gpuattract_2.0_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))


def load_package():
    if "gpuattract_amd" in sys.modules:
        return sys.modules["gpuattract_amd"]
    spec = importlib.util.spec_from_file_location(
        "gpuattract_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gpuattract_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def W():
    from gpuattract_amd import workload
    return workload


@pytest.fixture(scope="session")
def O():
    import oracle_py
    return oracle_py


@pytest.fixture(scope="session")
def wl_rigid(W):
    return W.make_workload("rigid small", 11, 400, 120, 300)


@pytest.fixture(scope="session")
def wl_modes(W):
    return W.make_workload("modes small", 12, 360, 110, 200, 3, 4)
