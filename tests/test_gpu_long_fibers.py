"""Fibers longer than the workgroup-local kernels take (meca.cc:323-444 factorises
any block; fiber_naked.cc:22-30 allows any length): blocks beyond LDS are
assembled column by column and inverted in device memory, fibers of up to 170
points keep the two-launch iteration, a longer one sends its system to the
four-launch iteration.  Blocks against the oracle's getBlock (1e-12) and its LU
inverse (FP64 1e-10, FP32 1e-4); positions after a step within 1e-9."""
import numpy as np
import pytest

from conftest import oracle_from_scene, rel_err
from test_gpu_parity import gpu_from_scene

pytestmark = pytest.mark.gpu

# fiber lengths (um, segmentation 0.5): 3D blocks of 153, 243, 333 and 483
# unknowns beside short fibers; the last scene adds a 241-point fiber
TILED = [25.0, 40.0, 55.0, 80.0] + [3.0, 4.5, 6.0, 2.5, 5.0, 3.5]
FOUR = [120.0, 30.0] + [3.0, 4.5, 6.0, 2.5, 5.0, 3.5]


def _scene(scenes, dim, lengths, seed):
    return scenes.mini(dim=dim, seed=seed, nlinks=40, radius=3.0, binding_range=0.6, lengths=lengths)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("precision", [1, 2])
def test_long_fiber_blocks(pkg, scenes, oracle, dim, precision):
    sc = _scene(scenes, dim, TILED, 41)
    print(f"long blocks dim={dim} precision={precision} points={sc.nb_points.tolist()}", flush=True)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    g = sc.gauss(0)
    gpu.setup(g)
    orc.setup(g)
    orc.compute_preconditioner()
    gpu.compute_preconditioner(precision)
    bar = 1e-10 if precision == pkg.PRECOND_FP64 else 1e-4
    for f in range(sc.nfib):
        assert rel_err(gpu.get_block(f), orc.get_block(f)) < 1e-12, f
        assert rel_err(gpu.get_block(f, inverse=True), orc.get_inverse_block(f)) < bar, f
    x = np.random.default_rng(5).standard_normal(sc.size)
    assert rel_err(gpu.precondition(x), orc.precondition(x)) < bar
    gpu.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("lengths", ["tiled", "four"])
@pytest.mark.parametrize("precondition", [0, 1])
def test_long_fiber_step(pkg, scenes, oracle, dim, lengths, precondition):
    sc = _scene(scenes, dim, TILED if lengths == "tiled" else FOUR, 43)
    print(f"long step dim={dim} {lengths} precondition={precondition} points={sc.nb_points.tolist()}", flush=True)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    g = sc.gauss(0)
    tol = 1e-12
    rep = gpu.solve(tol, g, tol_mode=pkg.TOL_RELATIVE, precondition=precondition)
    orc.solve(tol, g, tol_mode=1, precondition=precondition)
    assert rep["iterations"] > 0
    assert rel_err(gpu.fiber_points(), orc.fiber_points()) < 1e-9
    gpu.close()
