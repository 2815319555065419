"""Boundary behaviour of the C ABI that the reference's Meca fixes by its
calling sequence: apply needs a solve of the current setup, and per-fiber drag
settings (FiberProp::hydrodynamic_radius, surface_effect / cylinder_height,
fiber.cc:315-447) reach the mobility."""
import dataclasses

import numpy as np
import pytest

from conftest import oracle_assemble, rel_err

pytestmark = pytest.mark.gpu


def test_apply_needs_a_solve_of_the_current_setup(pkg, scenes, oracle):
    sc = scenes.mini(dim=3, seed=21, nfib=6, nlinks=15)
    gpu = pkg.Meca(3, 0)
    pkg.load_scene(gpu, sc)
    g = sc.gauss(0)
    gpu.setup(g)
    print("apply guard: apply after setup", flush=True)
    with pytest.raises(pkg.CymgError):
        gpu.apply()
    gpu.iterate(1e-11, tol_mode=pkg.TOL_RELATIVE)
    # the autotuner times launches on the solver's vectors: the solution is gone
    gpu.autotune(reps=2)
    with pytest.raises(pkg.CymgError):
        gpu.apply()
    gpu.iterate(1e-11, tol_mode=pkg.TOL_RELATIVE)
    gpu.time_kernel(6, reps=2)
    with pytest.raises(pkg.CymgError):
        gpu.apply()
    gpu.iterate(1e-11, tol_mode=pkg.TOL_RELATIVE)
    gpu.apply()
    from oracle import pyoracle
    orc = pyoracle.from_scene(sc)
    orc.solve(1e-11, g, tol_mode=1)
    assert rel_err(gpu.fiber_points(), orc.fiber_points()) < 1e-9
    gpu.close()


def test_per_fiber_drag_and_surface_model(pkg, scenes, oracle):
    """Fibers with their own hydrodynamic radii and, for a third of them, the
    surface drag model: one step matches the oracle given the same settings."""
    sc = scenes.mini(dim=3, seed=22, nfib=12, nlinks=30)
    r0 = np.where(np.arange(sc.nfib) % 2 == 0, 0.02, 0.0)
    r1 = np.where(np.arange(sc.nfib) % 4 == 1, 1.5, 0.0)
    height = np.where(np.arange(sc.nfib) % 3 == 2, 0.05, 0.0)
    gpu = pkg.Meca(3, 0)
    from oracle import pyoracle
    orc = pyoracle.OracleMeca(3, sc.params())
    for f in range(sc.nfib):
        gpu.add_fiber(sc.nb_points[f], sc.segmentation[f], sc.rigidity[f], r0[f], r1[f], height[f])
        orc.add_fiber(sc.nb_points[f], sc.segmentation[f], sc.rigidity[f], r0[f], r1[f], height[f])
    gpu.set_params(**sc.params())
    gpu.prepare(sc.points)
    pkg.assemble(gpu, sc)
    orc.prepare(sc.points)
    oracle_assemble(orc, sc)
    g = sc.gauss(0)
    print("per-fiber drag: one step", flush=True)
    gpu.solve(1e-11, g, tol_mode=pkg.TOL_RELATIVE)
    orc.solve(1e-11, g, tol_mode=1)
    assert rel_err(gpu.vector(pkg.VEC_RHS), orc.vector(oracle.RHS)) < 1e-12
    assert rel_err(gpu.fiber_points(), orc.fiber_points()) < 1e-9
    gpu.close()


def test_surface_drag_needs_a_height(pkg):
    gpu = pkg.Meca(3, 0)
    d = pkg.FiberDesc(5, 0.5, 20.0, 0.0, 0.0, 1, 0.0)
    import ctypes as C
    mi = C.c_uint32()
    assert gpu.L.cymg_add_fiber_ex(gpu.h, C.byref(d), C.byref(mi)) == pkg.ERR_INVALID
    assert b"must set and > 0" in gpu.L.cymg_last_error(gpu.h)
    gpu.close()


def test_dump_writes_the_reference_files(pkg, scenes, oracle, tmp_path):
    """Meca::dump (meca.cc:957-990): the system size, the drag per point of
    every Mecable, vRHS, vSOL, vPTS and the operator column by column."""
    sc = scenes.mini(dim=2, seed=5, nfib=4, nlinks=6)
    gpu = pkg.Meca(2, 0)
    pkg.load_scene(gpu, sc)
    g = sc.gauss(0)
    gpu.setup(g)
    gpu.iterate(1e-10, tol_mode=pkg.TOL_RELATIVE)
    print("dump: 2D mini scene", flush=True)
    gpu.dump(tmp_path)
    n = sc.size
    assert int((tmp_path / "ddim.txt").read_text()) == n
    for name, vec in (("drhs.bin", pkg.VEC_RHS), ("dsol.bin", pkg.VEC_SOL), ("dpts.bin", pkg.VEC_PTS)):
        assert np.array_equal(np.fromfile(tmp_path / name, dtype=np.float64), gpu.vector(vec))
    mat = np.fromfile(tmp_path / "dmat.bin", dtype=np.float64).reshape(n, n)   # row k = M e_k
    orc = oracle.from_scene(sc)
    orc.setup(g)
    dense = np.stack([orc.multiply(e) for e in np.eye(n)])
    assert rel_err(mat, dense) < 1e-12
    rows = np.loadtxt(tmp_path / "diagonal.txt")
    assert rows.shape == (n, 3)
    want_drag = np.concatenate([np.full(2 * int(sc.nb_points[f]), 1.0 / orc.mobility(f)) for f in range(sc.nfib)])
    assert np.array_equal(rows[:, 0].astype(int), np.repeat(np.arange(sc.nfib), 2 * sc.nb_points))
    assert rel_err(rows[:, 2], want_drag) < 1e-7   # printed with 9 significant digits
    assert rel_err(rows[:, 1], gpu.fiber_points().ravel()) < 1e-7
    gpu.close()
