import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(
    ROOT, "p-dit-probabilistic-ising-machine-for-solving-the-quadratic-assignment-problem_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_qpt():
    """Imports the product package (its directory name is not an identifier)."""
    if "qap_pt_amd" in sys.modules:
        return sys.modules["qap_pt_amd"]
    spec = importlib.util.spec_from_file_location(
        "qap_pt_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["qap_pt_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def qpt():
    return load_qpt()


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gpu(qpt):
    """The loaded HIP library; GPU tests fail (not skip) when it is missing."""
    lib = qpt.load_library()
    if qpt.device_count() < 1:
        pytest.fail("no HIP device visible to libqap_pt.so")
    return lib
