"""Parity at the benchmark workloads' full sizes: the 1000-fiber crosslinked
scene (assembly bit-exact, one implicit step), the aster scene (10 Solids of
201 points, 2000 fibers, 50000 motors; one step, then a motor step) and every
system of the 64-system batch, each against the oracle on the same inputs
at a tight tolerance."""
import dataclasses
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import oracle_from_scene, rel_err
from test_gpu_parity import _assert_assembly_equal, gpu_from_scene

pytestmark = pytest.mark.gpu


def test_crosslinked_1000_fibers_full_size(pkg, scenes, oracle):
    sc = scenes.config2(seed=2)
    print(f"full size: {sc.name} unknowns={sc.size}", flush=True)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    g = sc.gauss(0)
    gpu.setup(g)
    orc.setup(g)
    _assert_assembly_equal(pkg, oracle, gpu, orc)
    r1 = gpu.iterate(1e-10, tol_mode=pkg.TOL_RELATIVE)
    r2 = orc.iterate(1e-10, tol_mode=1)
    assert r1["converged"] and r2["converged"]
    gpu.apply()
    orc.apply()
    assert rel_err(gpu.fiber_points(), orc.fiber_points()) < 1e-9
    gpu.close()


def test_asters_full_size(pkg, scenes, oracle):
    sc = scenes.config3(seed=3)
    print(f"full size: {sc.name} unknowns={sc.size}", flush=True)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    g = sc.gauss(0)
    r1 = gpu.solve(1e-10, g, tol_mode=pkg.TOL_RELATIVE)
    r2 = orc.solve(1e-10, g, tol_mode=1)
    assert r1["converged"] and r2["converged"]
    assert rel_err(gpu.fiber_points(), orc.fiber_points()) < 1e-9
    gpu.step_couples(sc.motor_max_speed, sc.motor_stall_force)
    orc.step_couples(sc.link_fiber_a, sc.link_abs_a, sc.link_fiber_b, sc.link_abs_b, sc.link_stiffness,
                     sc.link_state, sc.motor_max_speed, sc.motor_stall_force)
    aa, ab, st = gpu.get_couples()
    assert np.array_equal(st, sc.link_state)
    assert np.abs(aa - sc.link_abs_a).max() < 1e-9 and np.abs(ab - sc.link_abs_b).max() < 1e-9
    gpu.close()


def test_batch_of_64_systems_full_size(pkg, scenes, oracle):
    print("full size: generating 64 systems", flush=True)
    batch = scenes.config5(seed=5)
    big = batch.scene
    print(f"full size: 64 systems, unknowns={big.size}", flush=True)
    gpu = pkg.Meca(3, 0)
    pkg.load_batch(gpu, batch)
    rep = gpu.solve(1e-12, big.gauss(0), tol_mode=pkg.TOL_RELATIVE, max_iter=6000)
    print(f"batch solved: converged={rep['converged']} iterations={rep['iterations']}", flush=True)
    assert rep["converged"]
    reps = gpu.system_reports()
    pts = gpu.fiber_points()
    gpu.close()
    offs = np.concatenate([[0], np.cumsum([sc.size for sc in batch.systems])])
    g = big.gauss(0)

    def oracle_step(k):
        sc = batch.systems[k]
        orc = oracle_from_scene(oracle, sc)
        orep = orc.solve(1e-12, g[offs[k]:offs[k + 1]], tol_mode=1)
        return k, orep["converged"], orc.fiber_points()

    # every system against its own oracle step; the oracle's solves run in
    # threads (one C call each, outside the interpreter lock), largest first
    order = sorted(range(batch.nsys), key=lambda k: -batch.systems[k].nfib)
    worst = 0.0
    with ThreadPoolExecutor(8) as ex:
        for k, conv, ref in ex.map(oracle_step, order):
            err = rel_err(pts[offs[k]:offs[k + 1]], ref)
            worst = max(worst, err)
            assert reps[k]["converged"] and conv, k
            assert err < 1e-9, (k, batch.systems[k].nfib, err)
    print(f"64 systems against the oracle: largest relative position error {worst:.2e}", flush=True)


def test_learned_fp32_on_the_1000_fiber_system(pkg, scenes, oracle):
    """The learned guess + FP32 preconditioner on the 1000-fiber crosslinked
    system, the network trained on the device from a rollout of plain FP64
    steps: a step of the running simulation solved at relative residual 1e-6
    lands within 1e-6 of the oracle's FP64 step (2-norm of the positions,
    relative), as the plain FP64 solve at the same residual does, whether the
    guard kept the network's guess or replaced it by zero (a guess whose
    max-norm residual is not below |b|_inf is dropped)."""
    sc = scenes.config2(seed=2)
    gpu = pkg.Meca(3, 0)
    pkg.load_scene(gpu, sc)
    print("learned fp32, 1000 fibers: rollout", flush=True)
    gpu.guess_record(40 * sc.nfib)
    for s in range(40):
        if s:
            gpu.prepare(None)
            pkg.assemble(gpu, sc, upload_couples=False)
        gpu.solve(1e-6, sc.gauss(500 + s), tol_mode=pkg.TOL_RELATIVE)
    gpu.guess_record(0)
    gpu.guess_init(3)
    gpu.guess_train(1500, batch=512, lr=1e-3, seed=4)
    start = gpu.fiber_points()
    print("learned fp32, 1000 fibers: step", flush=True)
    gpu.prepare(None)
    pkg.assemble(gpu, sc, upload_couples=False)
    g = sc.gauss(7)
    rep = gpu.solve(1e-6, g, tol_mode=pkg.TOL_RELATIVE, precondition=pkg.PRECOND_FP32, guess=pkg.GUESS_LEARNED)
    assert rep["converged"] and rep["guess_used"] in (pkg.GUESS_LEARNED, pkg.GUESS_ZERO)
    learned = gpu.fiber_points()
    # plain FP64 from the same points and draws
    plain = pkg.Meca(3, 0)
    pkg.load_scene(plain, dataclasses.replace(sc, points=start))
    rep64 = plain.solve(1e-6, g, tol_mode=pkg.TOL_RELATIVE)
    assert rep64["converged"]
    orc = oracle.OracleMeca(3, sc.params())
    oracle.register_scene(orc, sc)
    orc.prepare(start)
    oracle.assemble_scene(orc, sc)
    orc.solve(1e-12, g, tol_mode=1)
    ref = orc.fiber_points()
    e_learned = np.linalg.norm(learned - ref) / np.linalg.norm(ref)
    e_plain = np.linalg.norm(plain.fiber_points() - ref) / np.linalg.norm(ref)
    print(f"learned fp32 (guess used: {rep['guess_used']}) iterations {rep['iterations']} error {e_learned:.2e}; "
          f"plain fp64 iterations {rep64['iterations']} error {e_plain:.2e}", flush=True)
    assert e_learned < 1e-6 and e_plain < 1e-6
    gpu.close()
    plain.close()
