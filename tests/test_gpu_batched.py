"""Independent systems solved in one set of launches (cymg_set_systems): each
system gets the oracle's positions as if it were solved alone, with its own
Krylov scalars and stop test."""
import numpy as np
import pytest

from conftest import oracle_from_scene, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
def test_batch_matches_independent_solves(pkg, scenes, oracle, dim):
    batch = scenes.mini_batch(seed=9, nsys=4, dim=dim)
    gpu = pkg.Meca(dim, 0)
    pkg.load_batch(gpu, batch)
    big = batch.scene
    tol = 1e-11
    rep = gpu.solve(tol, big.gauss(0), tol_mode=pkg.TOL_RELATIVE)
    assert rep["converged"]
    reps = gpu.system_reports()
    pts = gpu.fiber_points()
    offs = [0]
    for sc in batch.systems:
        offs.append(offs[-1] + sc.size)
    for k, sc in enumerate(batch.systems):
        orc = oracle_from_scene(oracle, sc)
        orep = orc.solve(tol, sc.gauss(0), tol_mode=1)
        assert reps[k]["converged"] and orep["converged"]
        assert reps[k]["threshold"] == pytest.approx(orep["threshold"], rel=1e-12)
        assert rel_err(pts[offs[k]:offs[k + 1]], orc.fiber_points()) < 1e-9
    # systems stop independently
    assert len({r["iterations"] for r in reps}) >= 1


def test_reference_threshold_is_per_system(pkg, scenes, oracle):
    """noiseLevel (meca.cc:633-643) is the minimum over the system's own Mecables."""
    batch = scenes.mini_batch(seed=10, nsys=3, dim=3)
    gpu = pkg.Meca(3, 0)
    pkg.load_batch(gpu, batch)
    gpu.solve(0.05, batch.scene.gauss(0))
    reps = gpu.system_reports()
    for k, sc in enumerate(batch.systems):
        orc = oracle_from_scene(oracle, sc)
        orep = orc.solve(0.05, sc.gauss(0))
        assert reps[k]["noise_level"] == pytest.approx(orep["noise_level"], rel=1e-13)


def test_cross_system_coupling_rejected(pkg, scenes):
    batch = scenes.mini_batch(seed=11, nsys=2, dim=3)
    gpu = pkg.Meca(3, 0)
    pkg.load_batch(gpu, batch)
    last = int(batch.scene.nb_points.sum()) - 1
    with pytest.raises(pkg.CymgError):
        gpu.inter_link_ee(0, last, 10.0)
