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
    """Oracle Meca loaded with one step of the scene, in the call order of
    Simul::setInteractions: fibers' confinement, then couples."""
    m = pyoracle.OracleMeca(sc.dim, sc.params())
    for f in range(sc.nfib):
        m.add_fiber(sc.nb_points[f], sc.segmentation[f], sc.rigidity[f])
    m.prepare(sc.points)
    oracle_assemble(m, sc)
    return m


def oracle_assemble(m, sc):
    if sc.confine_radius > 0:
        m.confine_sphere(sc.confine_radius, sc.confine_stiffness)
    if sc.nlinks:
        a1, a2, ca = sc.interpolate(sc.link_fiber_a, sc.link_abs_a)
        b1, b2, cb = sc.interpolate(sc.link_fiber_b, sc.link_abs_b)
        for k in range(sc.nlinks):
            m.inter_link_ii(int(a1[k]), int(a2[k]), float(ca[k]), int(b1[k]), int(b2[k]),
                            float(cb[k]), float(sc.link_stiffness[k]))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
