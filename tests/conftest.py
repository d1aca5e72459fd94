"""Shared test setup.

Markers: `gpu` marks tests that need an MI355X and the built libhdg.
The oracle (oracle/) is the CPU checker; the product package lives in
hdiscontinuousgalerkin.jl_amd/ and is imported through `hdg_amd` below.
"""
import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built libhdg")


def load_package():
    """Import the product package (directory name contains a dot)."""
    if "hdg_amd" in sys.modules:
        return sys.modules["hdg_amd"]
    pkg_dir = os.path.join(REPO, "hdiscontinuousgalerkin.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "hdg_amd", os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hdg_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hdg():
    return load_package()


@pytest.fixture(scope="session")
def fixture_mesh_root():
    return os.path.join(GOLDEN, "figure2.1")
