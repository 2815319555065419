"""Solid Mecables (aster cores) on the device against the oracle: assembly of
the aster links bit-exact, setup vectors and operator within 1e-12, the
Solid's block (kept as identity + low rank on the device) equal to the
oracle's dense block and its inverse to the oracle's LU inverse, and fiber and
Solid positions over several steps (rescale, then a reshape on the 8th
getPoints) within 1e-9 relative at a tight tolerance."""
import numpy as np
import pytest

from conftest import oracle_assemble, oracle_from_scene, rel_err
from test_gpu_parity import gpu_from_scene, nr_entries

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("device_emitters", [True, False])
@pytest.mark.parametrize("short", [False, True])
def test_aster_assembly_bit_exact(pkg, scenes, oracle, dim, device_emitters, short):
    """With `short`, fibers shorter than the clamp distance are held by their
    plus end to the Solid's centre and clamp point interpolated at
    length/clampA (aster.cc:77-84, the device emitter's kind 2)."""
    sc = scenes.mini_asters(dim=dim, seed=42, lmin=0.3, lmax=1.4, core=0.8) if short else scenes.mini_asters(dim=dim)
    if short:
        assert (sc.aster_kind == 2).any()
    print(f"aster assembly dim={dim} device emitters={device_emitters} short={short}", flush=True)
    gpu = gpu_from_scene(pkg, sc, device_emitters)
    orc = oracle_from_scene(oracle, sc)
    g = sc.gauss(0)
    gpu.setup(g)
    orc.setup(g)
    for which in (0, 1):
        r, c, v = gpu.matrix_entries(which)
        rr, cc, vv = nr_entries(orc, which)
        keep = np.ones(rr.size, bool)
        touched = set(c.tolist())
        for i in range(rr.size):
            if rr[i] == cc[i] and cc[i] not in touched and vv[i] == 0.0:
                keep[i] = False
        rr, cc, vv = rr[keep], cc[keep], vv[keep]
        o1 = np.lexsort((r, c))
        o2 = np.lexsort((rr, cc))
        assert np.array_equal(r[o1], rr[o2]) and np.array_equal(c[o1], cc[o2])
        assert np.array_equal(v[o1], vv[o2])
    assert rel_err(gpu.vector(pkg.VEC_FOR), orc.vector(oracle.FOR)) < 1e-12
    assert rel_err(gpu.vector(pkg.VEC_RHS), orc.vector(oracle.RHS)) < 1e-12
    gpu.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_aster_operator_blocks_precondition(pkg, scenes, oracle, dim):
    sc = scenes.mini_asters(dim=dim, seed=12)
    print(f"aster operator dim={dim}", flush=True)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    g = sc.gauss(0)
    gpu.setup(g)
    orc.setup(g)
    rng = np.random.default_rng(5)
    for _ in range(3):
        x = rng.standard_normal(sc.size)
        assert rel_err(gpu.multiply(x), orc.multiply(x)) < 1e-12
    gpu.compute_preconditioner(pkg.PRECOND_FP64)
    orc.compute_preconditioner()
    for k in range(sc.nsol):
        m = sc.nfib + k
        B = gpu.get_block(m)
        assert rel_err(B, orc.get_block(m)) < 1e-12
        Bi = gpu.get_block(m, inverse=True)
        assert rel_err(Bi, orc.get_inverse_block(m)) < 1e-10
        assert np.abs(Bi @ B - np.eye(B.shape[0])).max() < 1e-10
    x = rng.standard_normal(sc.size)
    assert rel_err(gpu.precondition(x), orc.precondition(x)) < 1e-10
    gpu.close()


def _aster_steps(pkg, oracle, sc, steps, tol, precondition, guess=0, init_guess=False, launches=2, threads=512):
    gpu = gpu_from_scene(pkg, sc)
    gpu.set_option(pkg.OPT_KRYLOV_LAUNCHES, launches)
    gpu.set_option(pkg.OPT_TILE_THREADS, threads)
    assert gpu.get_option(pkg.OPT_TILE_THREADS) == threads
    if init_guess:
        gpu.guess_init(4)
        gpu.set_option(pkg.OPT_GUESS_GUARD, 0)
    orc = oracle_from_scene(oracle, sc)
    reps = []
    for s in range(steps):
        g = sc.gauss(s)
        print(f"aster step {s} precond={precondition} guess={guess}", flush=True)
        r1 = gpu.solve(tol, g, tol_mode=pkg.TOL_RELATIVE, precondition=precondition, guess=guess)
        r2 = orc.solve(tol, g, tol_mode=1, precondition=min(precondition, 1))
        reps.append((r1, r2))
        assert rel_err(gpu.fiber_points(), orc.fiber_points()) < (1e-9 if precondition != 2 else 1e-6), s
        if s + 1 < steps:
            gpu.prepare(None)
            pkg.assemble(gpu, sc, upload_couples=False)
            orc.prepare(orc.fiber_points())
            oracle_assemble(orc, sc)
    return gpu, orc, reps


@pytest.mark.parametrize("dim,launches,threads", [(2, 2, 512), (3, 2, 512), (3, 4, 512), (2, 2, 256), (3, 2, 256)])
def test_aster_steps_fp64(pkg, scenes, oracle, dim, launches, threads):
    """9 steps (every Solid gets at least one reshape: initial timers in
    [0, 7)) through the two-launch iteration (the Solids in tiles of their
    own, in workgroups of 512 or 256 threads) and the four-launch one."""
    sc = scenes.mini_asters(dim=dim, seed=13)
    gpu, orc, reps = _aster_steps(pkg, oracle, sc, 9, 1e-12, pkg.PRECOND_FP64, launches=launches, threads=threads)
    for r1, r2 in reps:
        assert r1["converged"] and r2["converged"]
    assert rel_err(gpu.vector(pkg.VEC_FOR), orc.vector(oracle.FOR)) < 1e-7
    assert rel_err(gpu.tensions(), orc.tensions()) < 1e-6
    gpu.close()


@pytest.mark.parametrize("precond", [0, 2])
@pytest.mark.parametrize("threads", [512, 256])
def test_aster_step_other_preconditioners(pkg, scenes, oracle, precond, threads):
    sc = scenes.mini_asters(dim=3, seed=14)
    gpu, orc, reps = _aster_steps(pkg, oracle, sc, 2, 1e-12 if precond == 0 else 1e-9, precond, threads=threads)
    assert all(r1["converged"] for r1, _ in reps)
    gpu.close()


def test_aster_step_learned_guess(pkg, scenes, oracle):
    sc = scenes.mini_asters(dim=3, seed=15)
    gpu, orc, reps = _aster_steps(pkg, oracle, sc, 2, 1e-12, pkg.PRECOND_FP64, guess=pkg.GUESS_LEARNED,
                                  init_guess=True)
    assert all(r1["guess_used"] == pkg.GUESS_LEARNED for r1, _ in reps)
    gpu.close()


@pytest.mark.parametrize("threads", [512, 256])
def test_asters_reduced_config3(pkg, scenes, oracle, threads):
    """The aster scene's shape (Solids of 1 + N points, fibers of 11-21 points,
    motors) at 3 asters x 40 fibers."""
    sc = scenes.asters(dim=3, seed=16, nasters=3, fibers_per_aster=40, nmotors=1500, spread=2.0, radius=8.0)
    gpu, orc, reps = _aster_steps(pkg, oracle, sc, 2, 1e-11, pkg.PRECOND_FP64, threads=threads)
    assert all(r1["converged"] for r1, _ in reps)
    gpu.close()


def test_solid_registration_rules(pkg):
    m = pkg.Meca(3, 0)
    m.add_fiber(5, 0.5, 20.0)
    m.add_solid([0.5, 0.0, 0.0], np.arange(9, dtype=float))
    with pytest.raises(pkg.CymgError):
        m.add_fiber(5, 0.5, 20.0)          # fibers come first
    with pytest.raises(pkg.CymgError):
        m.set_systems([0, 1])               # no batches with Solids
    m.close()


@pytest.mark.parametrize("precondition", [1, 2, 0])
def test_aster_graph_replay_matches_eager(pkg, scenes, precondition):
    """Solid tiles in the graph-replayed iteration chunks give the same bits as
    the same launches issued one by one."""
    sc = scenes.mini_asters(dim=3, seed=17)
    out = []
    for graphs in (1, 0):
        gpu = gpu_from_scene(pkg, sc)
        gpu.set_option(pkg.OPT_GRAPHS, graphs)
        its = []
        for s in range(3):
            print(f"aster graph replay precond={precondition} graphs={graphs} step={s}", flush=True)
            if s:
                gpu.prepare(None)
                pkg.assemble(gpu, sc, upload_couples=False)
            its.append(gpu.solve(1e-8, sc.gauss(s), tol_mode=pkg.TOL_RELATIVE, precondition=precondition)["iterations"])
        out.append((gpu.fiber_points(), its))
        gpu.close()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
