"""Registration order changing between steps, as NodeList::mix reshuffles the
Mecables every step (node_list.cc:244-280, simul_step.cc:344-350) and objects
come and go: one context re-registers the fibers in a new order (and with one
fiber fewer) each step, from the current points, and every step matches the
oracle built in the same order (positions within 1e-9 relative).  This runs
the layout change through the tiles, the captured iteration graphs, the
block offsets and the reused index structures of the assembly."""
import dataclasses

import numpy as np
import pytest

from conftest import oracle_from_scene, rel_err
from test_gpu_parity import gpu_from_scene

pytestmark = pytest.mark.gpu


def _reordered(sc, order, points):
    """The scene with fibers `order` (a permutation of a subset) registered in
    that order, at `points` (current points in sc's order); couples touching a
    dropped fiber are dropped."""
    D = sc.dim
    starts = np.concatenate([[0], np.cumsum(sc.nb_points)]).astype(np.int64)
    P = np.asarray(points).reshape(-1, D)
    pts = np.concatenate([P[starts[f]:starts[f + 1]] for f in order])
    new_index = np.full(sc.nfib, -1, np.int64)
    new_index[order] = np.arange(order.size)
    keep = (new_index[sc.link_fiber_a] >= 0) & (new_index[sc.link_fiber_b] >= 0)
    return dataclasses.replace(
        sc, nb_points=sc.nb_points[order].astype(np.uint32), segmentation=sc.segmentation[order],
        rigidity=sc.rigidity[order], points=pts.ravel(),
        link_fiber_a=new_index[sc.link_fiber_a[keep]].astype(np.uint32), link_abs_a=sc.link_abs_a[keep],
        link_fiber_b=new_index[sc.link_fiber_b[keep]].astype(np.uint32), link_abs_b=sc.link_abs_b[keep],
        link_stiffness=sc.link_stiffness[keep],
        link_state=None if sc.link_state is None else sc.link_state[keep])


@pytest.mark.parametrize("dim", [2, 3])
def test_registration_order_changes_every_step(pkg, scenes, oracle, dim):
    sc = scenes.config2(seed=31 + dim, nfib=60, nlinks=300, radius=4.0) if dim == 3 else \
        scenes.mini(dim=2, seed=33, nfib=20, nlinks=60, radius=2.5)
    gpu = gpu_from_scene(pkg, sc)
    rng = np.random.default_rng(dim)
    for step in range(4):
        print(f"reorder: dim={dim} step={step} fibers={sc.nfib} links={sc.nlinks}", flush=True)
        orc = oracle_from_scene(oracle, sc)
        g = sc.gauss(step)
        r1 = gpu.solve(1e-11, g, tol_mode=pkg.TOL_RELATIVE)
        r2 = orc.solve(1e-11, g, tol_mode=1)
        assert r1["converged"] and r2["converged"]
        assert rel_err(gpu.fiber_points(), orc.fiber_points()) < 1e-9, step
        # next step: a new order, one fiber fewer on the third step, from the oracle's points
        order = rng.permutation(sc.nfib)
        if step == 2:
            order = order[:-1]
        sc = _reordered(sc, order, orc.fiber_points())
        pkg.load_scene(gpu, sc)
    gpu.close()


def test_identities_carry_points_and_previous_solution(pkg, scenes, oracle):
    """The reference's calling pattern with a warm start: every step clear(),
    add() in a new order (one fiber gone on the second step), identities given,
    prepare(NULL) from the device's own points.  The device must hand each
    fiber its own points and its share of the previous solution (bit for bit),
    and the step seeded from it must match the oracle."""
    sc = scenes.config2(seed=41, nfib=50, nlinks=250, radius=4.0)
    ids = np.arange(1000, 1000 + sc.nfib, dtype=np.uint32)
    gpu = pkg.Meca(3, 0)
    pkg.load_scene(gpu, sc)
    gpu.set_identities(ids)
    rng = np.random.default_rng(5)
    orc = oracle_from_scene(oracle, sc)
    for step in range(3):
        print(f"identities: step={step} fibers={sc.nfib}", flush=True)
        g = sc.gauss(step)
        r1 = gpu.solve(1e-11, g, tol_mode=pkg.TOL_RELATIVE, guess=pkg.GUESS_PREVIOUS if step else pkg.GUESS_ZERO)
        r2 = orc.solve(1e-11, g, tol_mode=1)
        assert r1["converged"] and r2["converged"]
        assert r1["guess_used"] == (pkg.GUESS_PREVIOUS if step else pkg.GUESS_ZERO)
        pts = gpu.fiber_points()
        assert rel_err(pts, orc.fiber_points()) < 1e-9, step
        sol = gpu.vector(pkg.VEC_SOL)
        order = rng.permutation(sc.nfib)
        if step == 0:
            order = order[:-1]
        starts = np.concatenate([[0], np.cumsum(sc.nb_points)]).astype(np.int64)
        sc = _reordered(sc, order, pts)
        ids = ids[order]
        expect_prev = np.concatenate([sol[3 * starts[f]:3 * starts[f + 1]] for f in order])
        # the GPU context: new order from its own points
        gpu.clear()
        for f in range(sc.nfib):
            gpu.add_fiber(sc.nb_points[f], sc.segmentation[f], sc.rigidity[f])
        gpu.set_identities(ids)
        gpu.prepare(None)
        pkg.assemble(gpu, sc)
        assert np.array_equal(gpu.vector(pkg.VEC_PREV), expect_prev)
        assert np.array_equal(gpu.vector(pkg.VEC_PTS), sc.points)
        orc = oracle_from_scene(oracle, sc)
    gpu.close()


def test_without_identities_a_new_order_forgets(pkg, scenes):
    """Re-registered in a new order without identities: the previous solution
    is dropped (zero) and prepare(NULL) refuses, since the device points no
    longer match the Mecables."""
    sc = scenes.config2(seed=42, nfib=20, nlinks=60, radius=3.0)
    gpu = pkg.Meca(3, 0)
    pkg.load_scene(gpu, sc)
    gpu.solve(1e-8, sc.gauss(0), tol_mode=pkg.TOL_RELATIVE)
    # same registration again: kept
    gpu.clear()
    for f in range(sc.nfib):
        gpu.add_fiber(sc.nb_points[f], sc.segmentation[f], sc.rigidity[f])
    gpu.prepare(None)
    assert np.abs(gpu.vector(pkg.VEC_PREV)).max() > 0
    order = np.roll(np.arange(sc.nfib), 3)
    sc2 = _reordered(sc, order, gpu.fiber_points())
    gpu.clear()
    for f in range(sc2.nfib):
        gpu.add_fiber(sc2.nb_points[f], sc2.segmentation[f], sc2.rigidity[f])
    with pytest.raises(pkg.CymgError):
        gpu.prepare(None)
    gpu.prepare(sc2.points)
    assert not np.any(gpu.vector(pkg.VEC_PREV))
    gpu.close()
