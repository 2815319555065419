"""Pins the oracle's steric interactions (Simul::setStericInteractions,
simul_solve.cc:139-188, with PointGrid, point_grid.cc) before the device grid is
compared with it.

* The grid finds every interaction: its calls equal, as a multiset, those of the
  same checks run over every pair of items (brute force), on random scenes of
  fibers and Solids in 2D and 3D.
* Sphere-sphere contacts (Solid points) are exactly the pairs closer than the
  sum of their radii, computed here with numpy.
* Two crossing fibers closer than twice their steric radius are pushed apart
  along the line of closest approach with force w (2r - d).
* A steric_max_range too short for the segments is refused (point_grid.cc:86-95).
* The calls come in the grid's cell order: the cell of each sphere contact's
  first sphere, from the reference's grid dimensions, never decreases.
"""
import numpy as np
import pytest

PARAMS = dict(time_step=0.005, kT=0.0042, viscosity=0.02)


def _walk(rng, dim, n, cut, box):
    start = rng.uniform(-box, box, dim)
    d = rng.standard_normal(dim)
    d /= np.linalg.norm(d)
    pts = [start]
    for _ in range(n - 1):
        d = d + 0.3 * rng.standard_normal(dim)
        d /= np.linalg.norm(d)
        pts.append(pts[-1] + cut * d)
    return np.concatenate(pts)


def _scene(oracle, dim, seed, nfib=40, nsol=6, box=1.6, cut=0.3):
    rng = np.random.default_rng(seed)
    m = oracle.OracleMeca(dim, PARAMS)
    pts = []
    for f in range(nfib):
        n = int(rng.integers(3, 9))
        m.add_fiber(n, cut, 20.0)
        pts.append(_walk(rng, dim, n, cut, box))
    for k in range(nsol):
        n = int(rng.integers(3, 7))
        rad = rng.uniform(0.03, 0.15, n)
        rad[rng.uniform(size=n) < 0.2] = 0.0
        p = rng.uniform(-box, box, dim) + 0.2 * rng.standard_normal((n, dim))
        m.add_solid(rad, p.ravel())
        pts.append(p.ravel())
    m.prepare(np.concatenate(pts))
    return m, nfib, nsol


def _flags(nfib, nsol, radius=0.05, extra=0.05):
    n = nfib + nsol
    st = np.ones(n, np.int32)
    rad = np.full(n, radius)
    rng = np.full(n, extra)
    return st, rad, rng


def _key(call):
    kind, p, ca, cb, ln, w = call
    return (kind, p, ca, cb, ln, w)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_grid_finds_every_pair_the_brute_force_finds(oracle, dim, seed):
    m, nfib, nsol = _scene(oracle, dim, seed)
    st, rad, rng = _flags(nfib, nsol)
    ext = [2.5] * dim
    grid = m.steric(ext, 0.0, 300.0, 100.0, st, rad, rng)
    m2, _, _ = _scene(oracle, dim, seed)
    brute = m2.steric(ext, 0.0, 300.0, 100.0, st, rad, rng, brute_force=True)
    assert len(grid) > 20
    assert sorted(map(_key, grid)) == sorted(map(_key, brute))
    kinds = {c[0] for c in grid}
    assert kinds >= {0, 1} and (dim == 2 or 2 in kinds)


@pytest.mark.parametrize("dim", [2, 3])
def test_sphere_contacts_are_the_pairs_closer_than_their_radii(oracle, dim):
    rng = np.random.default_rng(10 + dim)
    m = oracle.OracleMeca(dim, PARAMS)
    pts, rads, idx = [], [], []
    base = 0
    for k in range(8):
        n = int(rng.integers(3, 8))
        rad = rng.uniform(0.05, 0.2, n)
        rad[0] = 0.0   # no steric sphere for a point of zero radius
        p = rng.uniform(-1.0, 1.0, dim) + 0.3 * rng.standard_normal((n, dim))
        m.add_solid(rad, p.ravel())
        pts.append(p)
        rads.append(rad)
        idx.append(np.arange(base, base + n))
        base += n
    m.prepare(np.concatenate([p.ravel() for p in pts]))
    P = np.concatenate(pts)
    R = np.concatenate(rads)
    I = np.concatenate(idx)
    n = len(pts)
    calls = m.steric([1.8] * dim, 0.0, 300.0, 100.0, np.ones(n, np.int32), np.zeros(n), np.zeros(n))
    got = {tuple(sorted(c[1][:2])) for c in calls}
    assert all(c[0] == 0 for c in calls)
    want = set()
    live = np.nonzero(R > 128 * np.finfo(float).eps)[0]
    for a in live:
        for b in live:
            if a < b and np.linalg.norm(P[a] - P[b]) < R[a] + R[b]:
                want.add((int(I[a]), int(I[b])))
    assert got == want and len(want) > 5


def test_crossing_fibers_pushed_apart(oracle):
    """Two perpendicular one-segment fibers crossing near their middles at
    distance d < 2r in 3D: the segment-segment check makes one
    interSideSlidingLink (point_grid.cc:342-351) and no other check fires (the
    ends are far from the other segment); the recorded force on each fiber is
    w (2r - d) along the common normal, away from the other."""
    m = oracle.OracleMeca(3, PARAMS)
    cut, r, d, w = 0.6, 0.05, 0.06, 300.0
    a = np.array([[-0.3, 0, 0], [0.3, 0, 0]], float) + [0.01, 0.02, 0]
    b = np.array([[0, -0.3, d], [0, 0.3, d]], float) + [0.05, 0.04, 0]
    m.add_fiber(2, cut, 20.0)
    m.add_fiber(2, cut, 20.0)
    x0 = np.concatenate([a.ravel(), b.ravel()])
    m.prepare(x0)
    calls = m.steric([1.0] * 3, 0.0, w, 50.0, np.ones(2, np.int32), np.full(2, r), np.full(2, 0.02))
    assert [c[0] for c in calls] == [2]
    B = m.matrix_dense(0)
    C = m.matrix_dense(1)
    m.setup(None)
    F = np.kron(B, np.eye(3)) @ x0 + C @ x0 + m.vector(oracle.BAS)
    fa = F[:6].reshape(2, 3).sum(0)
    fb = F[6:].reshape(2, 3).sum(0)
    assert np.allclose(fa, [0, 0, -w * (2 * r - d)], atol=1e-9 * w)
    assert np.allclose(fb, -fa, atol=1e-9 * w)


def test_range_too_short_is_refused(oracle):
    m, nfib, nsol = _scene(oracle, 3, 4)
    st, rad, rng = _flags(nfib, nsol)
    with pytest.raises(ValueError):
        m.steric([2.5] * 3, 0.2, 300.0, 100.0, st, rad, rng)


def _effective(calls):
    """the calls that contribute: Meca skips a link between points that share a
    model point (PointInterpolated::overlapping, point_exact.cc neighbours)"""
    out = []
    for c in calls:
        kind, p = c[0], c[1]
        if kind == 1 and p[2] in (p[0], p[1]):
            continue
        if kind == 2 and {p[0], p[1]} & {p[2], p[3]}:
            continue
        out.append(c)
    return out


def _force(oracle, m, dim, x0):
    B = m.matrix_dense(0)
    C = m.matrix_dense(1)
    m.setup(None)
    return np.kron(B, np.eye(dim)) @ x0 + C @ x0 + m.vector(oracle.BAS)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("where", ["middle", "beyond_plus", "before_minus", "outside_joint"])
def test_sphere_against_a_fiber(oracle, dim, where):
    """checkPL (point_grid.cc:125-172) on a three-point fiber bent by 60 degrees
    at its middle point: a Solid point beside the middle of a segment gets an
    interSideSlidingLink pushing it away along the normal by w (r1 + r2 - d);
    beyond the plus end, before the minus end, or outside the joint (before the
    second segment, past the first), it meets that fiber point as two spheres
    (interLongLink)."""
    m = oracle.OracleMeca(dim, PARAMS)
    cut, rf, rs, w = 0.6, 0.05, 0.1, 300.0
    fib = np.zeros((3, dim))
    fib[0, 0] = -cut
    fib[2, :2] = [cut * 0.5, cut * np.sqrt(0.75)]
    q = {"middle": [-0.25, 0.08], "beyond_plus": [0.34, 0.56], "before_minus": [-0.66, 0.02],
         "outside_joint": [0.05, -0.05]}[where]
    sol = np.zeros((2, dim))
    sol[0, :2] = q
    sol[1, :2] = [0.0, -1.5]
    m.add_fiber(3, cut, 20.0)
    m.add_solid(np.array([rs, 0.0]), sol.ravel())
    x0 = np.concatenate([fib.ravel(), sol.ravel()])
    m.prepare(x0)
    calls = _effective(m.steric([2.0] * dim, 0.0, w, 50.0, np.ones(2, np.int32), np.array([rf, 0.0]),
                                np.array([0.02, 0.0])))
    F = _force(oracle, m, dim, x0)
    fq = F[3 * dim:4 * dim]
    if where == "middle":
        assert [c[0] for c in calls] == [1]
        want = np.zeros(dim)
        want[1] = w * (rf + rs - q[1])
        assert np.allclose(fq, want, atol=1e-9 * w)
    else:
        end = {"beyond_plus": 2, "before_minus": 0, "outside_joint": 1}[where]
        assert [(c[0], c[1][0], c[1][1]) for c in calls] == [(0, 3, end)]
        v = sol[0] - fib[end]
        dist = np.linalg.norm(v)
        assert np.allclose(fq, w * (rf + rs - dist) * v / dist, atol=1e-9 * w)
    assert np.allclose(F.reshape(-1, dim).sum(0), 0, atol=1e-9 * w)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("end", ["minus", "plus"])
@pytest.mark.parametrize("gap", [0.07, 0.12])
@pytest.mark.parametrize("first", ["segment", "tip"])
def test_fiber_end_against_a_segment(oracle, dim, end, gap, first):
    """checkLL1 / checkLL2 (point_grid.cc:178-315): the minus (or plus) end of a
    three-point fiber beside the middle of another fiber's first segment gets an
    interSideSlidingLink and no other check fires; closer than 2r it is pushed
    away with the push stiffness (w (2r - d)), within r + r + range it is pulled
    back with the pull stiffness.  Either fiber may come first in the cells
    (registration order), which decides the branch of checkLL that sees it."""
    m = oracle.OracleMeca(dim, PARAMS)
    cut, r, push, pull = 0.6, 0.05, 300.0, 40.0
    a = np.zeros((3, dim))
    a[:, 0] = [-0.3, 0.3, 0.9]
    b = np.zeros((3, dim))
    b[:, 0] = 0.05
    b[:, 1] = [gap, gap + cut, gap + 2 * cut] if end == "minus" else [gap + 2 * cut, gap + cut, gap]
    m.add_fiber(3, cut, 20.0)
    m.add_fiber(3, cut, 20.0)
    order = [a, b] if first == "segment" else [b, a]
    ia = 0 if first == "segment" else 3                    # first point of the segment's fiber
    ib = 3 - ia                                            # first point of the tip's fiber
    x0 = np.concatenate([o.ravel() for o in order])
    m.prepare(x0)
    calls = _effective(m.steric([2.0] * dim, 0.0, push, pull, np.ones(2, np.int32), np.full(2, r),
                                np.full(2, 0.03)))
    tip = ib + (0 if end == "minus" else 2)
    w = push if gap < 2 * r else pull
    assert [(c[0], c[1][0], c[1][2], c[5]) for c in calls] == [(1, ia, tip, w)]
    assert abs(calls[0][2] - 0.35 / cut) < 1e-12
    F = _force(oracle, m, dim, x0)
    want = np.zeros(dim)
    want[1] = w * (2 * r - gap)
    assert np.allclose(F[tip * dim:tip * dim + dim], want, atol=1e-9 * push)
    assert np.allclose(F.reshape(-1, dim).sum(0), 0, atol=1e-9 * push)


def _cells(pos, dim, ext, step):
    """Cell index of each position in the reference's PointGrid: per axis
    ceil(2 ext / step) + 2 cells of width `step` centred on the origin
    (point_grid.cc:40-50), positions clamped to the edge cells, x fastest
    (Grid::index, grid.h:445-453)."""
    n = [int(np.ceil(2 * ext / step)) + 2 for _ in range(dim)]
    inf = [-0.5 * n[d] * step for d in range(dim)]
    idx = np.zeros(len(pos), np.int64)
    ijk = np.zeros((len(pos), dim), np.int64)
    for d in reversed(range(dim)):
        c = np.clip(np.floor((pos[:, d] - inf[d]) / step).astype(np.int64), 0, n[d] - 1)
        ijk[:, d] = c
        idx = idx * n[d] + c
    return idx, ijk


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", [4, 5])
def test_sphere_calls_follow_the_cell_order(oracle, dim, seed):
    """The order of the calls (the order in which the reference assembles the
    contributions): the grid visits its cells in index order, so the cell of
    each contact's first sphere, computed here from the grid's dimensions,
    never decreases along the calls, and the second sphere is in that cell or
    a neighbour.  A grid without its two border cells makes the same calls in
    another order and fails this."""
    rng = np.random.default_rng(seed)
    m = oracle.OracleMeca(dim, PARAMS)
    pts = []
    for k in range(60):
        n = int(rng.integers(4, 9))
        rad = rng.uniform(0.05, 0.14, n)
        box = 1.9 if dim == 2 else 1.2
        p = rng.uniform(-box, box, dim) + 0.25 * rng.standard_normal((n, dim))
        m.add_solid(rad, p.ravel())
        pts.append(p)
    pos = np.concatenate(pts)
    m.prepare(pos.ravel())
    st = np.ones(60, np.int32)
    ext, step = 2.0, 0.3
    calls = m.steric([ext] * dim, step, 300.0, 100.0, st, np.zeros(60), np.zeros(60))
    assert len(calls) > 100 and all(c[0] == 0 for c in calls)
    first = np.array([c[1][0] for c in calls])
    second = np.array([c[1][1] for c in calls])
    ca, ija = _cells(pos[first], dim, ext, step)
    cb, ijb = _cells(pos[second], dim, ext, step)
    assert np.all(np.diff(ca) >= 0), "contacts out of the grid's cell order"
    assert np.abs(ija - ijb).max() <= 1
