"""Pins of the oracle's motor couples (orc_step_couples: Couple::stepAA/AF/FA,
couple.cc:125-150, Motor::stepLoaded/stepUnloaded, motor.cc:19-57) by what
the motor model fixes: an unloaded motor walks at its unloaded speed, a motor
held back by its stall force stops, one pushed forward by the stall force runs
at twice its speed (the limit_speed cap), a larger backward load cannot make it
walk backwards unless limit_speed is off, passive hands never move, a hand that
walks past a fiber end holds the end or detaches, and a motor whose partner
hand is free ignores the spring."""
import numpy as np
import pytest

from oracle import pyoracle as po

PRM = dict(time_step=0.005, kT=0.0042, viscosity=0.02, hydro_radius0=0.0125, hydro_radius1=10.0)
DT = PRM["time_step"]


def _two_fibers(dim, gap=0.2, npts=9):
    """Two straight fibers along +x (segments of exactly 0.5), the second
    shifted by `gap` along y."""
    m = po.OracleMeca(dim, PRM)
    m.add_fiber(npts, 0.5, 20.0)
    m.add_fiber(npts, 0.5, 20.0)
    p = np.zeros((2 * npts, dim))
    p[:npts, 0] = 0.5 * np.arange(npts)
    p[npts:, 0] = 0.5 * np.arange(npts)
    p[npts:, 1] = gap
    m.prepare(p.ravel())
    return m


def _step(m, aa, ab, w, state, speed=1.0, stall=5.0, limit=True, hold=False):
    aa = np.array(aa, dtype=np.float64)
    ab = np.array(ab, dtype=np.float64)
    st = np.array(state, dtype=np.uint8)
    m.step_couples(np.zeros(aa.size, np.uint32), aa, np.ones(aa.size, np.uint32), ab,
                   np.asarray(w, dtype=np.float64), st, speed, stall, limit, hold)
    return aa, ab, st


@pytest.mark.parametrize("dim", [2, 3])
def test_unloaded_motor_walks_at_its_speed(dim):
    m = _two_fibers(dim)
    aa, ab, st = [1.0, 2.0], [1.0, 2.0], [15, 7]
    for k in range(1, 6):
        aa, ab, st = _step(m, aa, ab, [0.0, 0.0], st, speed=2.0)
        assert np.allclose(aa, [1.0 + 2.0 * DT * k, 2.0 + 2.0 * DT * k], rtol=0, atol=1e-14)
        assert np.allclose(ab, [1.0 + 2.0 * DT * k, 2.0], rtol=0, atol=1e-14)   # hand 2 of couple 1 is passive
    assert list(st) == [15, 7]


@pytest.mark.parametrize("dim", [2, 3])
def test_stall_force_stops_and_assisting_load_doubles_speed(dim):
    stall, w = 5.0, 100.0
    m = _two_fibers(dim)
    d = stall / w        # hand 2 behind hand 1 by d along x: the spring pulls hand 1 back with the stall force
    aa, ab, st = _step(m, [2.0], [2.0 - d], [w], [15], speed=1.0, stall=stall)
    assert abs(aa[0] - 2.0) < 1e-14                       # held back by its stall force: no motion
    assert abs(ab[0] - (2.0 - d) - 2 * DT) < 1e-14        # pulled forward by the stall force: twice the speed


@pytest.mark.parametrize("dim", [2, 3])
def test_speed_limits(dim):
    stall, w = 5.0, 100.0
    m = _two_fibers(dim)
    d = 3 * stall / w    # three times the stall force
    aa, ab, st = _step(m, [2.0], [2.0 - d], [w], [15], stall=stall)
    assert aa[0] == 2.0                                   # backward drive capped at 0
    assert abs(ab[0] - (2.0 - d) - 2 * DT) < 1e-14        # forward drive capped at 2 v dt
    aa, ab, st = _step(m, [2.0], [2.0 - d], [w], [15], stall=stall, limit=False)
    assert abs(aa[0] - (2.0 - 2 * DT)) < 1e-14            # v dt - 3 v dt without the cap
    assert abs(ab[0] - (2.0 - d) - 4 * DT) < 1e-14


def test_minus_end_directed_motor():
    m = _two_fibers(3)
    aa, ab, st = _step(m, [2.0], [3.0], [0.0], [15], speed=-1.0)
    assert abs(aa[0] - (2.0 - DT)) < 1e-15 and abs(ab[0] - (3.0 - DT)) < 1e-15


def test_passive_hands_never_move():
    m = _two_fibers(3)
    aa, ab, st = _step(m, [1.0, 2.5], [3.0, 0.7], [100.0, 50.0], [3, 1], speed=3.0)
    assert list(aa) == [1.0, 2.5] and list(ab) == [3.0, 0.7] and list(st) == [3, 1]


@pytest.mark.parametrize("hold", [False, True])
def test_fiber_end(hold):
    m = _two_fibers(3, npts=5)                            # length 2.0
    aa, ab, st = _step(m, [2.0 - 0.5 * DT], [0.3], [0.0], [15], speed=1.0, hold=hold)
    if hold:
        assert aa[0] == 2.0 and st[0] == 15               # moveToEnd(PLUS_END)
    else:
        assert st[0] == 14                                # hand 1 detached
    assert abs(ab[0] - (0.3 + DT)) < 1e-15
    # a detached couple's remaining motor walks unloaded; the spring no longer acts
    if not hold:
        aa2, ab2, st2 = _step(m, aa, ab, [1e6], st, speed=1.0)
        assert abs(ab2[0] - (0.3 + 2 * DT)) < 1e-15 and aa2[0] == aa[0] and st2[0] == 14


def test_free_partner_ignores_spring():
    m = _two_fibers(3)
    aa, ab, st = _step(m, [1.0], [5.0], [1e6], [13], speed=1.0)   # hand 1 attached motor, hand 2 free
    assert abs(aa[0] - (1.0 + DT)) < 1e-15 and ab[0] == 5.0 and st[0] == 13
