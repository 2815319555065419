"""Motor couples stepped on the device between implicit steps (Couple::stepAA
/ stepAF / stepFA with Motor::stepLoaded / stepUnloaded) against the oracle:
hand abscissas within 1e-9 (absolute, µm) and attachment states equal after
every step, with hands crossing segments and falling off (or holding) fiber
ends, and the following implicit steps (assembled from the moved couples on
the device) within 1e-9 relative of the oracle's."""
import numpy as np
import pytest

from conftest import oracle_assemble, oracle_from_scene, rel_err
from test_gpu_parity import gpu_from_scene

pytestmark = pytest.mark.gpu


def _motor_scene(scenes, dim, seed):
    sc = scenes.mini(dim=dim, seed=seed, nfib=10, nlinks=40, radius=1.6)
    rng = np.random.default_rng(seed)
    sc.link_state = rng.choice(np.array([3, 7, 11, 15, 15, 13, 14], np.uint8), sc.nlinks)
    sc.motor_max_speed = 40.0     # 0.2 um per step: hands cross segments and reach the ends
    sc.motor_stall_force = 5.0
    return sc


@pytest.mark.parametrize("dim,hold", [(3, False), (2, False), (3, True)])
def test_motor_steps_match_oracle(pkg, scenes, oracle, dim, hold):
    sc = _motor_scene(scenes, dim, 31 + dim)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    st0 = sc.link_state.copy()
    for s in range(6):
        print(f"motor step dim={dim} hold={hold} step={s}", flush=True)
        g = sc.gauss(s)
        r1 = gpu.solve(1e-12, g, tol_mode=pkg.TOL_RELATIVE)
        r2 = orc.solve(1e-12, g, tol_mode=1)
        assert r1["converged"] and r2["converged"]
        assert rel_err(gpu.fiber_points(), orc.fiber_points()) < 1e-9, s
        gpu.step_couples(sc.motor_max_speed, sc.motor_stall_force, hold_growing_end=hold)
        orc.step_couples(sc.link_fiber_a, sc.link_abs_a, sc.link_fiber_b, sc.link_abs_b, sc.link_stiffness,
                         sc.link_state, sc.motor_max_speed, sc.motor_stall_force, hold_growing_end=hold)
        aa, ab, st = gpu.get_couples()
        assert np.array_equal(st, sc.link_state), s
        assert np.abs(aa - sc.link_abs_a).max() < 1e-9 and np.abs(ab - sc.link_abs_b).max() < 1e-9, s
        gpu.prepare(None)
        pkg.assemble(gpu, sc, upload_couples=False)
        orc.prepare(orc.fiber_points())
        oracle_assemble(orc, sc)
    moved = (np.abs(sc.link_abs_a - scenes.mini(dim=dim, seed=31 + dim, nfib=10, nlinks=40,
                                                 radius=1.6).link_abs_a) > 0.5).sum()
    assert moved > 0                                        # hands crossed segments
    if not hold:
        assert (sc.link_state != st0).any()                 # some hands fell off an end
    gpu.close()


def test_couple_state_validation(pkg, scenes):
    sc = scenes.mini(dim=3, seed=3, nfib=4, nlinks=5)
    gpu = gpu_from_scene(pkg, sc)
    with pytest.raises(pkg.CymgError):
        gpu.set_couple_state(np.full(sc.nlinks, 16, np.uint8))
    with pytest.raises(pkg.CymgError):
        gpu.step_couples(1.0, 0.0)                          # motor:stall_force must be > 0
    gpu.close()
