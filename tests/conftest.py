import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def scenes(pkg):
    import importlib
    return importlib.import_module(__graft_entry__.PKG_NAME + ".scene")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    return pyoracle


def oracle_from_scene(pyoracle, sc):
    """Oracle Meca loaded with one step of the scene."""
    return pyoracle.from_scene(sc)


def oracle_assemble(m, sc):
    from oracle import pyoracle
    pyoracle.assemble_scene(m, sc)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
