"""Staged launches of the two-launch iteration (CYMG_OPT_BLOCK_STAGE): four
more wavefronts per 512-thread tile copy the tile's block inverses into LDS
for the block apply (value 2: FP64 and FP32 blocks; 1, the default: FP32
blocks only).  The staged columns are summed in the same order as the
columns read from memory, so the solve must give the same bits as the
unstaged launches, and match the oracle like them.  Tiles whose blocks exceed
the 128 KB stage read the remaining columns from memory."""
import numpy as np
import pytest

from conftest import oracle_from_scene, rel_err
from test_gpu_parity import gpu_from_scene

pytestmark = pytest.mark.gpu


def _steps(pkg, sc, stage, precondition, tol, steps=2, tile_points=0):
    gpu = gpu_from_scene(pkg, sc)
    gpu.set_option(pkg.OPT_BLOCK_STAGE, stage)
    gpu.set_option(pkg.OPT_TILE_POINTS, tile_points)
    assert gpu.get_option(pkg.OPT_BLOCK_STAGE) == stage
    reps = []
    for s in range(steps):
        reps.append(gpu.solve(tol, sc.gauss(s), tol_mode=1, precondition=precondition))
        if s + 1 < steps:
            gpu.prepare(None)
            pkg.assemble(gpu, sc, upload_couples=False)
    out = (gpu.fiber_points().copy(), gpu.tensions().copy(), [r["iterations"] for r in reps])
    gpu.close()
    return out, reps


def _same_bits(a, b):
    assert a[2] == b[2], f"iterations {a[2]} != {b[2]}"
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("precondition", [1, 2])
def test_staged_launches_give_the_same_bits(pkg, scenes, oracle, dim, precondition):
    """Fibers of 2-15 um (tiles of a few fibers, most within the stage), the
    FP64 and FP32 blocks, 2D and 3D: the same points, tensions and
    iterations as the unstaged launches, and the oracle's points."""
    sc = scenes.config2(seed=41 + dim, nfib=200, nlinks=1500, radius=6.0) if dim == 3 else \
        scenes.mini(dim=2, seed=41, nfib=60, nlinks=200, radius=4.0, lmin=2.0, lmax=15.0)
    print(f"staged launches: dim={dim} precondition={precondition}", flush=True)
    tol = 1e-11 if precondition == 1 else 1e-9
    ref, _ = _steps(pkg, sc, 0, precondition, tol)
    got, reps = _steps(pkg, sc, 2, precondition, tol)
    _same_bits(ref, got)
    for r in reps:
        assert r["converged"]
    orc = oracle_from_scene(oracle, sc)
    orc.solve(tol, sc.gauss(0), tol_mode=1)
    gpu1, _ = _steps(pkg, sc, 2, precondition, tol, steps=1)
    assert rel_err(gpu1[0], orc.fiber_points()) < (1e-9 if precondition == 1 else 1e-6)


@pytest.mark.parametrize("precondition,tile_points", [(1, 0), (1, 170), (2, 170)])
def test_staged_launches_beyond_the_stage(pkg, scenes, precondition, tile_points):
    """Fibers of 29-31 points: two per 85-point tile are 138 KB of FP64 blocks
    (more than the stage holds), and five per 170-point tile (one thread per
    unknown) 345 KB FP64 or 173 KB FP32: the columns past the stage come
    from memory, with the same bits as the unstaged launches."""
    sc = scenes.config2(seed=47, nfib=60, nlinks=400, radius=6.0, lmin=14.0, lmax=15.0)
    assert sc.nb_points.min() >= 29
    print(f"staged launches beyond the stage: precondition={precondition} tile_points={tile_points}", flush=True)
    tol = 1e-11 if precondition == 1 else 1e-9
    ref, _ = _steps(pkg, sc, 0, precondition, tol, tile_points=tile_points)
    got, reps = _steps(pkg, sc, 2, precondition, tol, tile_points=tile_points)
    _same_bits(ref, got)
    for r in reps:
        assert r["converged"]


def test_staged_launch_time(pkg, scenes):
    """The staged full-work launch is measurable; a solve without a block
    preconditioner runs the unstaged launches (the option then has no effect)."""
    sc = scenes.config2(seed=48, nfib=200, nlinks=1500, radius=6.0)
    gpu = gpu_from_scene(pkg, sc)
    assert gpu.get_option(pkg.OPT_BLOCK_STAGE) == 1
    gpu.set_option(pkg.OPT_BLOCK_STAGE, 2)
    r = gpu.solve(1e-9, sc.gauss(0), tol_mode=1, precondition=0)
    assert r["converged"]
    r = gpu.solve(1e-9, sc.gauss(0), tol_mode=1, precondition=1)
    assert r["converged"]
    staged = gpu.time_kernel(6, reps=4)
    gpu.set_option(pkg.OPT_BLOCK_STAGE, 0)
    plain = gpu.time_kernel(6, reps=4)
    print(f"full-work launch: staged {1e3 * staged:.2f} us, unstaged {1e3 * plain:.2f} us", flush=True)
    assert staged > 0 and plain > 0
    gpu.close()


def test_staged_fp32_full_size_same_bits(pkg, scenes):
    """The 1000-fiber crosslinked system (240 tiles), one step with FP32 blocks
    at relative residual 1e-6: the default (staged) launches and the unstaged
    ones give the same points, tensions and iterations."""
    sc = scenes.config2(seed=2)
    print("staged FP32, 1000 fibers", flush=True)
    ref, _ = _steps(pkg, sc, 0, 2, 1e-6, steps=1)
    got, reps = _steps(pkg, sc, 1, 2, 1e-6, steps=1)
    assert reps[0]["converged"]
    _same_bits(ref, got)


def test_autotuned_staging_same_bits(pkg, scenes):
    """cymg_autotune with FP32 blocks chooses the tile shape and whether to
    stage them (1) or not (0); a solve on its choice gives the bits of the
    unstaged launches on the same tiles."""
    sc = scenes.config2(seed=49, nfib=200, nlinks=1500, radius=6.0)
    gpu = gpu_from_scene(pkg, sc)
    gpu.setup(sc.gauss(0))
    gpu.compute_preconditioner(pkg.PRECOND_FP32)
    t = gpu.autotune(reps=6)
    print(f"autotune with FP32 blocks: {t}", flush=True)
    assert t["block_stage"] in (0, 1) and t["launch_ms"] > 0
    r = gpu.solve(1e-9, sc.gauss(0), tol_mode=1, precondition=2)
    assert r["converged"]
    tuned = gpu.fiber_points().copy()
    gpu.close()
    ref = gpu_from_scene(pkg, sc)
    ref.set_option(pkg.OPT_TILE_THREADS, t["tile_threads"])
    ref.set_option(pkg.OPT_TILE_POINTS, t["tile_points"])
    ref.set_option(pkg.OPT_BLOCK_INVERSE, t["block_inverse"])
    ref.set_option(pkg.OPT_BLOCK_STAGE, 0)
    r0 = ref.solve(1e-9, sc.gauss(0), tol_mode=1, precondition=2)
    assert r0["iterations"] == r["iterations"]
    assert np.array_equal(ref.fiber_points(), tuned)
    ref.close()
