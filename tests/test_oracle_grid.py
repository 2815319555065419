"""Pins of the oracle's segment query (orc_nearby_segments: FiberGrid::
nearbySegments, fiber_grid.cc:302-337, with FiberLocus::projectPoint,
fiber_locus.cc:62-97) by what the geometry fixes: distances to a straight
fiber in closed form, the ends of a fiber seen only by its first and last
segments, a place projecting beyond an inner segment listed by its
neighbour only, the strict and inclusive range tests at a tie, the excluded
fiber, the abscissa from the minus end, and, on random fibers with segments
of exactly the segmentation, the kept set equal to the segments whose true
distance (computed independently) is below the range."""
import numpy as np
import pytest

from oracle import pyoracle as po

PRM = dict(time_step=0.005, kT=0.0042, viscosity=0.02, hydro_radius0=0.0125, hydro_radius1=10.0)


def _straight(dim, npts=5, cut=0.5, offsets=((0.0, 0.0),)):
    m = po.OracleMeca(dim, PRM)
    pts = []
    for off in offsets:
        m.add_fiber(npts, cut, 20.0)
        p = np.zeros((npts, dim))
        p[:, 0] = off[0] + cut * np.arange(npts)
        p[:, 1] = off[1]
        pts.append(p)
    m.prepare(np.concatenate(pts).ravel())
    return m


@pytest.mark.parametrize("dim", [2, 3])
def test_closed_form_distances(dim):
    m = _straight(dim)
    w = np.zeros(dim)
    w[0], w[1] = 1.1, 0.25
    f, p, a, d = m.nearby_segments(w, 0.3)
    assert list(f) == [0] and list(p) == [2]                  # inside segment 2 (x in [1.0, 1.5])
    assert abs(a[0] - 1.1) < 1e-15 and abs(d[0] - 0.0625) < 1e-15
    assert len(m.nearby_segments(w, 0.2)[0]) == 0


def test_fiber_ends_and_inner_segments():
    m = _straight(3)
    # before the minus end: only segment 0 measures the distance to its first point
    f, p, a, d = m.nearby_segments([-0.1, 0.1, 0.0], 0.5)
    assert list(p) == [0] and abs(d[0] - 0.02) < 1e-15 and abs(a[0] + 0.1) < 1e-15
    # after the plus end (x = 2.0): only the last segment
    f, p, a, d = m.nearby_segments([2.2, 0.0, 0.0], 0.5)
    assert list(p) == [3] and abs(d[0] - 0.04) < 1e-14
    # on the line of the fiber between points: the segment holding the projection and, when the
    # projection lands exactly on a shared point, both segments
    f, p, a, d = m.nearby_segments([1.0, 0.1, 0.0], 0.5)
    assert list(p) == [1, 2]
    assert np.allclose(a, [1.0, 1.0], rtol=0, atol=1e-15) and np.allclose(d, [0.01, 0.01], rtol=0, atol=1e-15)


def test_strict_and_inclusive_range_at_a_tie():
    m = _straight(3)
    w = [1.25, 0.25, 0.0]        # distance 0.25 exactly: dis = 0.0625 = range^2
    assert len(m.nearby_segments(w, 0.25, inclusive=False)[0]) == 0      # nearbySegments: dis < DD
    assert list(m.nearby_segments(w, 0.25, inclusive=True)[1]) == [2]    # tryToAttach: dis > range^2 rejected


def test_exclude_and_fiber_order():
    m = _straight(3, offsets=((0.0, 0.0), (0.0, 0.2), (0.0, -0.2)))
    f, p, a, d = m.nearby_segments([0.75, 0.0, 0.0], 0.3)
    assert list(f) == [0, 1, 2] and list(p) == [1, 1, 1]
    f, p, a, d = m.nearby_segments([0.75, 0.0, 0.0], 0.3, exclude=1)
    assert list(f) == [0, 2]


def _true_dist(w, P, Q):
    d = Q - P
    t = np.clip(np.dot(w - P, d) / np.dot(d, d), 0.0, 1.0)
    return np.linalg.norm(w - (P + t * d))


@pytest.mark.parametrize("dim", [2, 3])
def test_random_fibers_against_true_distance(dim):
    rng = np.random.default_rng(dim)
    cut, npts, nfib = 0.4, 7, 30
    m = po.OracleMeca(dim, PRM)
    pts = []
    for _ in range(nfib):
        m.add_fiber(npts, cut, 20.0)
        # a random walk of steps of exactly `cut` (bent fibers)
        p = np.zeros((npts, dim))
        p[0] = rng.uniform(-2, 2, dim)
        u = rng.standard_normal(dim)
        for k in range(1, npts):
            u = u + 0.6 * rng.standard_normal(dim)
            p[k] = p[k - 1] + cut * u / np.linalg.norm(u)
        pts.append(p)
    P = np.concatenate(pts)
    m.prepare(P.ravel())
    R = 0.35
    for _ in range(200):
        w = rng.uniform(-3, 3, dim)
        f, p, a, d = m.nearby_segments(w, R)
        got = set(zip(f.tolist(), p.tolist()))
        for fi in range(nfib):
            for s in range(npts - 1):
                A, B = P[fi * npts + s], P[fi * npts + s + 1]
                t = np.dot(w - A, B - A) / cut ** 2
                inner = 0 <= t <= 1
                end_ok = (t < 0 and s == 0) or (t > 1 and s == npts - 2)
                dist = _true_dist(w, A, B)
                if abs(dist - R) < 1e-9:
                    continue                                        # tie: either answer is rounding
                if inner or end_ok:
                    assert ((fi, s) in got) == (dist < R), (fi, s, dist)
                else:
                    assert (fi, s) not in got                       # projection beyond an inner segment
        for (fi, s), dd in zip(zip(f.tolist(), p.tolist()), d):
            A, B = P[fi * npts + s], P[fi * npts + s + 1]
            assert abs(np.sqrt(max(dd, 0.0)) - _true_dist(w, A, B)) < 1e-9
