"""Pins of the oracle's Solid Mecable (oracle/meca_oracle.c, solid.cc):
the mobility reduces forces to the exact inverse of the bead resistance of a
rigid body (independently derived here), the operator's diagonal block and its
inverse, the reshape that recovers a known rotation (Horn's quaternion
method) and the rescale that restores the reference size."""
import numpy as np
import pytest

from oracle import pyoracle as po

PRM = dict(time_step=0.005, kT=0.0042, viscosity=0.02, hydro_radius0=0.0125, hydro_radius1=10.0)


def _solid(dim, seed, n=9, timer=0, with_fiber=False):
    rng = np.random.default_rng(seed)
    m = po.OracleMeca(dim, PRM)
    pts = []
    if with_fiber:
        m.add_fiber(4, 0.5, 20.0)
        fp = np.zeros((4, dim))
        fp[:, 0] = 0.5 * np.arange(4)
        pts.append(fp.ravel())
    radius = rng.uniform(0.05, 0.4, n)
    radius[rng.integers(0, n, n // 3)] = 0.0          # drag-free points
    radius[0] = 0.3
    sp = rng.standard_normal((n, dim)) * 0.7 + rng.standard_normal(dim)
    i = m.add_solid(radius, sp.ravel(), timer)
    pts.append(sp.ravel())
    allp = np.concatenate(pts)
    m.prepare(allp)
    return m, i, radius, sp, allp


def _resistance(dim, radius, x, t, w, eta):
    """Total force and torque about the origin of beads of drag 6 pi eta r
    moving rigidly with (t, w), plus the rotational drag 8 pi eta r^3 w."""
    xi = 6 * np.pi * eta * radius
    zeta = 8 * np.pi * eta * (radius ** 3).sum()
    if dim == 2:
        v = t[None, :] + w * np.stack([-x[:, 1], x[:, 0]], axis=1)
        F = (xi[:, None] * v).sum(0)
        N = (xi * (x[:, 0] * v[:, 1] - x[:, 1] * v[:, 0])).sum() + zeta * w
        return F, np.array([N])
    v = t[None, :] + np.cross(w[None, :], x)
    F = (xi[:, None] * v).sum(0)
    N = (xi[:, None] * np.cross(x, v)).sum(0) + zeta * w
    return F, N


def _forces_with_moments(dim, x, F, N):
    """A force per point whose sum is F and moment about the origin is N."""
    n = x.shape[0]
    k = dim + (1 if dim == 2 else 3)
    rows = []
    for p in range(n):
        for d in range(dim):
            e = np.zeros(dim)
            e[d] = 1
            mom = [x[p, 0] * e[1] - x[p, 1] * e[0]] if dim == 2 else list(np.cross(x[p], e))
            rows.append(list(e) + mom)
    A = np.array(rows).T                      # k x (dim n)
    X, *_ = np.linalg.lstsq(A, np.concatenate([F, N]), rcond=None)
    assert A.shape[0] == k
    return X


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mobility_inverts_bead_resistance(dim, seed):
    m, i, radius, x, allp = _solid(dim, seed)
    rng = np.random.default_rng(seed + 10)
    t = rng.standard_normal(dim)
    w = rng.standard_normal() if dim == 2 else rng.standard_normal(3)
    F, N = _resistance(dim, radius, x, t, w, PRM["viscosity"])
    X = _forces_with_moments(dim, x, F, N)
    # drive the Solid's setSpeedsFromForces through the RHS of a setup:
    # vRHS = dt P vFOR, with vFOR = the pure forces X
    m2 = po.OracleMeca(dim, PRM)
    m2.add_solid(radius, x.ravel(), 0)
    m2.prepare(x.ravel())
    for p in range(x.shape[0]):
        f = np.zeros(3)
        f[:dim] = X[dim * p:dim * p + dim]
        m2.add_pure_force(p, f)
    m2.setup(None)
    rhs = m2.vector(po.RHS).reshape(-1, dim)
    if dim == 2:
        want = t[None, :] + w * np.stack([-x[:, 1], x[:, 0]], axis=1)
    else:
        want = t[None, :] + np.cross(w[None, :], x)
    want = want * PRM["time_step"]            # RHS = dt * P * forces
    assert np.allclose(rhs, want, rtol=0, atol=1e-11 * np.abs(want).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_solid_drag_and_centre(dim):
    m, i, radius, x, allp = _solid(dim, 4)
    s = m.solid(i)
    eta = PRM["viscosity"]
    assert np.isclose(s["drag"], 6 * np.pi * eta * radius.sum(), rtol=1e-14)
    assert np.isclose(s["drag_rot"], 8 * np.pi * eta * (radius ** 3).sum(), rtol=1e-14)
    c = (radius[:, None] * x).sum(0) / radius.sum()
    assert np.allclose(s["center"][:dim], c, rtol=0, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_solid_block_is_operator_block(dim):
    # getBlock equals the operator's diagonal block (the DEBUG_MECA==4 check,
    # meca.cc:335-347), and the stored inverse inverts it
    m, i, radius, x, allp = _solid(dim, 5, n=7, with_fiber=True)
    rng = np.random.default_rng(3)
    n_pts = m.nb_points
    for k in range(6):
        a, b = rng.integers(0, n_pts, 2)
        if a != b:
            m.inter_link_ee(int(a), int(b), 30.0)
    m.setup(None)
    o = m.dim * int(m.mat_index(i))
    bs = m.dim * x.shape[0]
    blk = m.get_block(i)
    cols = np.zeros((bs, bs))
    for j in range(bs):
        e = np.zeros(m.size)
        e[o + j] = 1.0
        cols[:, j] = m.multiply(e)[o:o + bs]
    assert np.allclose(blk, cols, rtol=0, atol=1e-13)
    m.compute_preconditioner()
    inv = m.get_inverse_block(i)
    assert np.allclose(inv @ blk, np.eye(bs), rtol=0, atol=1e-10)


def _rotation(dim, rng):
    if dim == 2:
        a = rng.uniform(0, 2 * np.pi)
        return np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


@pytest.mark.parametrize("dim", [2, 3])
def test_reshape_recovers_rotation(dim):
    rng = np.random.default_rng(dim)
    n = 12
    radius = np.zeros(n)
    radius[0] = 0.5
    shape = rng.standard_normal((n, dim))
    m = po.OracleMeca(dim, PRM)
    i = m.add_solid(radius, shape.ravel(), 7)         # the next getPoints reshapes
    Rm = _rotation(dim, rng)
    t = rng.standard_normal(dim)
    moved = shape @ Rm.T + t
    # a small deformation the reshape removes
    noisy = moved + 1e-3 * rng.standard_normal(moved.shape)
    m.prepare(noisy.ravel())
    m.setup(None)
    m.iterate(1e-10, tol_mode=1, precondition=0)
    m.apply()
    got = m.fiber_points().reshape(n, dim)
    cc = noisy.mean(0)
    want = (shape - shape.mean(0)) @ Rm.T + cc
    assert np.allclose(got, want, rtol=0, atol=5e-3)
    assert m.solid_timer(i) == 0
    # exactly rigid input: the reshape is exact
    m2 = po.OracleMeca(dim, PRM)
    m2.add_solid(radius, shape.ravel(), 7)
    m2.prepare(moved.ravel())
    m2.setup(None)
    m2.iterate(1e-10, tol_mode=1, precondition=0)
    m2.apply()
    assert np.allclose(m2.fiber_points().reshape(n, dim), moved, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_rescale_restores_size(dim):
    rng = np.random.default_rng(7 + dim)
    n = 10
    radius = np.zeros(n)
    radius[0] = 0.5
    shape = rng.standard_normal((n, dim))
    m = po.OracleMeca(dim, PRM)
    i = m.add_solid(radius, shape.ravel(), 0)         # the next getPoints rescales
    grown = 1.01 * (shape - shape.mean(0)) + shape.mean(0) + 0.2
    m.prepare(grown.ravel())
    m.setup(None)
    m.iterate(1e-10, tol_mode=1, precondition=0)
    m.apply()
    got = m.fiber_points().reshape(n, dim)
    assert np.allclose(got.mean(0), grown.mean(0), rtol=0, atol=1e-13)
    size = ((got - got.mean(0)) ** 2).sum()
    ref = ((shape - shape.mean(0)) ** 2).sum()
    assert np.isclose(size, ref, rtol=1e-12)
    assert m.solid_timer(i) == 1


@pytest.mark.parametrize("dim", [2, 3])
def test_aster_step_converges_and_keeps_solid_rigid(dim, scenes):
    sc = scenes.mini_asters(dim=dim)
    m = po.from_scene(sc)
    rep = m.solve(1e-10, sc.gauss(0), tol_mode=1)
    assert rep["converged"]
    pts = m.fiber_points()
    for k in range(sc.nsol):
        i = int(sc.sol_index[k])
        before = sc.solid_points(k).reshape(-1, dim)
        after = pts[dim * i:dim * (i + before.shape[0])].reshape(-1, dim)
        # the linearised rigid motion changes distances at second order; the
        # rescale restores the size exactly, the reshape (every 8th step) the shape
        d0 = np.linalg.norm(before[:, None] - before[None], axis=2)
        d1 = np.linalg.norm(after[:, None] - after[None], axis=2)
        assert np.allclose(d0, d1, rtol=0, atol=1e-2 * d0.max())
        size0 = ((before - before.mean(0)) ** 2).sum()
        size1 = ((after - after.mean(0)) ** 2).sum()
        if m.solid_timer(sc.nfib + k) == 0:
            assert np.allclose(d0, d1, rtol=0, atol=1e-10)
        else:
            assert np.isclose(size1, size0, rtol=1e-12)
    # noise level: the minimum of dt * b / drag (Solid) and dt * mobility * b (fibers)
    assert rep["noise_level"] > 0
