"""Pins the oracle's position-dependent emitters on interpolated points
(meca_inter.cc) against their force laws.

Each emitter records a linearised force  F(x) = vBAS + (mB (x) I + mC) x.
For every emitter whose force is Hookean once its geometry (fiber direction,
side arm) is frozen at the current positions x0, that force is -grad E of an
energy written here from the geometry alone (side points built with rotations
and cross products, never from the emitter's coefficients), and the recorded
F equals -grad E (finite differences) at x0 and at displaced positions.  The
non-linear ones (long links, torque) are compared at x0 with -grad E and, where
the reference linearises exactly, their matrix with the finite-difference
Hessian.  Side points are also checked geometrically at x0: at distance len
from the attachment, orthogonal to the fiber, on the side of the target.
"""
import numpy as np
import pytest

W = 37.5


def _chain(dim, np_, cut, rng, offset):
    d = 0.4 * rng.standard_normal((np_ - 1, dim)) + np.eye(dim)[0]
    d = cut * d / np.linalg.norm(d, axis=1, keepdims=True)
    return (np.vstack([np.zeros(dim), np.cumsum(d, axis=0)]) + offset).ravel()


def _meca(oracle, dim, cut=0.5, seed=1):
    """two fibers: points 0-4 and 5-8"""
    rng = np.random.default_rng(seed)
    pts = np.concatenate([_chain(dim, 5, cut, rng, 0.0),
                          _chain(dim, 4, cut, rng, 0.35 * np.ones(dim) + 0.6 * np.eye(dim)[1])])
    m = oracle.OracleMeca(dim, dict(time_step=0.005, kT=0.0042, viscosity=0.02))
    m.add_fiber(5, cut, 20.0)
    m.add_fiber(4, cut, 20.0)
    m.prepare(pts)
    return m, pts


def _recorded(oracle, m, dim):
    B = m.matrix_dense(0)
    C = m.matrix_dense(1)
    m.setup(None)
    return np.kron(B, np.eye(dim)) + C, m.vector(oracle.BAS)


def _pt(x, dim, i):
    return x[dim * i:dim * i + dim]


def _interp(x, dim, p1, p2, c):
    return (1 - c) * _pt(x, dim, p1) + c * _pt(x, dim, p2)


def _r90(v):
    return np.array([-v[1], v[0]])


def _cross2(a, b):
    return a[0] * b[1] - a[1] * b[0]


def _grad(E, x, h=1e-4):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e[i] = h
        g[i] = (E(x + e) - E(x - e)) / (2 * h)
    return g


def _hess(E, x, h=1e-4):
    n = x.size
    H = np.zeros((n, n))
    I = np.eye(n) * h
    for i in range(n):
        for j in range(i, n):
            H[i, j] = H[j, i] = (E(x + I[i] + I[j]) - E(x + I[i] - I[j]) - E(x - I[i] + I[j])
                                 + E(x - I[i] - I[j])) / (4 * h * h)
    return H


def _check_quadratic(M, bas, E, x0, rng, tol=1e-7):
    """F = -grad E at x0 and at two displaced configurations"""
    for x in (x0, x0 + 0.3 * rng.standard_normal(x0.size), x0 + rng.standard_normal(x0.size)):
        F = M @ x + bas
        G = -_grad(E, x)
        assert np.abs(F - G).max() < tol * W * max(1.0, np.abs(G).max() / W), (F, G)


def _side_frame(dim, d0, v0, length):
    """side point offset at x0 for a fiber of direction d0 and a target at v0
    from the attachment: orthogonal to the fiber, towards the target"""
    if dim == 2:
        s = 1.0 if _cross2(d0, v0) >= 0 else -1.0
        off = s * length * _r90(d0) / np.linalg.norm(d0)
        return s, off
    a = d0 / np.linalg.norm(d0)
    u = v0 - (v0 @ a) * a
    off = length * u / np.linalg.norm(u)
    return 1.0, off


def _assert_side_point(off, d0, v0, length):
    assert abs(np.linalg.norm(off) - length) < 1e-12 * length
    assert abs(off @ d0) < 1e-12 * np.linalg.norm(d0) * length
    assert off @ v0 > 0


A1, A2, CA = 1, 2, 0.3
B, B1, B2, CB = 6, 6, 7, 0.65


def test_pure_force_interpolated(oracle):
    """addPureForce(PointInterpolated) (meca_inter.cc:140-156): the force is
    shared by the two points with the interpolation weights"""
    for dim in (2, 3):
        m, pts = _meca(oracle, dim)
        f = np.arange(1.0, dim + 1.0)
        m.add_pure_force_i(A1, A2, CA, f)
        M, bas = _recorded(oracle, m, dim)
        assert not M.any()
        want = np.zeros(pts.size)
        want[dim * A1:dim * A1 + dim] = (1 - CA) * f
        want[dim * A2:dim * A2 + dim] = CA * f
        assert np.array_equal(bas, want)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("target", ["exact", "interpolated"])
@pytest.mark.parametrize("stretch", [1.6, 0.6])
def test_long_links(oracle, dim, target, stretch):
    """interLongLink PI-PE (meca_inter.cc:425-505) and PI-PI (516-596):
    E = w/2 (|B - A| - len)^2.  The recorded force is -grad E at x0; stretched,
    the matrix is the Hessian; compressed, it is -w (c c') (x) (u u') with c the
    interpolation coefficients and u the link direction (meca_inter.cc:460-462)."""
    m, x0 = _meca(oracle, dim)

    def ends(x):
        A = _interp(x, dim, A1, A2, CA)
        Bp = _pt(x, dim, B) if target == "exact" else _interp(x, dim, B1, B2, CB)
        return A, Bp

    A0, B0 = ends(x0)
    length = np.linalg.norm(B0 - A0) / stretch
    if target == "exact":
        m.inter_long_link_ie(A1, A2, CA, B, length, W)
        idx, c = [A1, A2, B], np.array([1 - CA, CA, -1.0])
    else:
        m.inter_long_link_ii(A1, A2, CA, B1, B2, CB, length, W)
        idx, c = [A1, A2, B1, B2], np.array([1 - CA, CA, -(1 - CB), -CB])
    M, bas = _recorded(oracle, m, dim)

    def E(x):
        A, Bp = ends(x)
        return 0.5 * W * (np.linalg.norm(Bp - A) - length) ** 2

    G = -_grad(E, x0, 1e-6)
    assert np.abs(M @ x0 + bas - G).max() < 1e-7 * W
    if stretch > 1:
        assert np.abs(M + _hess(E, x0)).max() < 1e-4 * W
    else:
        u = (B0 - A0) / np.linalg.norm(B0 - A0)
        want = np.zeros_like(M)
        blk = -W * np.kron(np.outer(c, c), np.outer(u, u))
        dofs = np.concatenate([np.arange(dim * i, dim * i + dim) for i in idx])
        want[np.ix_(dofs, dofs)] = blk
        assert np.abs(M - want).max() < 1e-12 * W


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("target", ["exact", "interpolated"])
def test_side_links(oracle, dim, target):
    """interSideLink PI-PE (meca_inter.cc:617-759) and PI-PI (764-935): a link
    of zero resting length between the target and S = A + arm, arm of length
    len orthogonal to the fiber towards the target.  2D: E = w/2 |B - S|^2 with
    S = A + e R90 (P2 - P1), e frozen.  3D: the force is projected on the plane
    of the fiber and the arm: E = w/2 |T (B - A - arm)|^2, T = a a' + b b'."""
    rng = np.random.default_rng(dim)
    m, x0 = _meca(oracle, dim)

    def tgt(x):
        return _pt(x, dim, B) if target == "exact" else _interp(x, dim, B1, B2, CB)

    d0 = _pt(x0, dim, A2) - _pt(x0, dim, A1)
    A0 = _interp(x0, dim, A1, A2, CA)
    length = 0.2
    s, off = _side_frame(dim, d0, tgt(x0) - A0, length)
    _assert_side_point(off, d0, tgt(x0) - A0, length)
    if target == "exact":
        m.inter_side_link_ie(A1, A2, CA, B, length, W)
    else:
        m.inter_side_link_ii(A1, A2, CA, B1, B2, CB, length, W)
    M, bas = _recorded(oracle, m, dim)
    if dim == 2:
        e = s * length / np.linalg.norm(d0)

        def E(x):
            S = _interp(x, dim, A1, A2, CA) + e * _r90(_pt(x, dim, A2) - _pt(x, dim, A1))
            return 0.5 * W * np.sum((tgt(x) - S) ** 2)
    else:
        a = d0 / np.linalg.norm(d0)
        b = off / length
        T = np.outer(a, a) + np.outer(b, b)

        def E(x):
            v = T @ (tgt(x) - _interp(x, dim, A1, A2, CA) - off)
            return 0.5 * W * v @ v
    _check_quadratic(M, bas, E, x0, rng)


@pytest.mark.parametrize("dim", [2, 3])
def test_side_side_link(oracle, dim):
    """interSideSideLink (meca_inter.cc:943-1059): a link between SA = A +
    len/2 N_A and SB = B + len/2 N_B, each side point orthogonal to its fiber
    towards the other fiber; E = w/2 |SA - SB|^2.  No 3D version (it throws)."""
    m, x0 = _meca(oracle, dim)
    length = 0.3
    if dim == 3:
        with pytest.raises(ValueError):
            m.inter_side_side_link(A1, A2, CA, B1, B2, CB, length, W)
        return
    rng = np.random.default_rng(5)
    da = _pt(x0, 2, A2) - _pt(x0, 2, A1)
    db = _pt(x0, 2, B2) - _pt(x0, 2, B1)
    A0 = _interp(x0, 2, A1, A2, CA)
    B0 = _interp(x0, 2, B1, B2, CB)
    s1, off1 = _side_frame(2, da, B0 - A0, length / 2)
    s2, off2 = _side_frame(2, db, A0 - B0, length / 2)
    _assert_side_point(off1, da, B0 - A0, length / 2)
    _assert_side_point(off2, db, A0 - B0, length / 2)
    m.inter_side_side_link(A1, A2, CA, B1, B2, CB, length, W)
    M, bas = _recorded(oracle, m, 2)
    e1 = s1 * length / 2 / np.linalg.norm(da)
    e2 = s2 * length / 2 / np.linalg.norm(db)

    def E(x):
        SA = _interp(x, 2, A1, A2, CA) + e1 * _r90(_pt(x, 2, A2) - _pt(x, 2, A1))
        SB = _interp(x, 2, B1, B2, CB) + e2 * _r90(_pt(x, 2, B2) - _pt(x, 2, B1))
        return 0.5 * W * np.sum((SA - SB) ** 2)
    _check_quadratic(M, bas, E, x0, rng)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("target", ["exact", "interpolated"])
def test_sliding_links(oracle, dim, target):
    """interSlidingLink PI-PE (meca_inter.cc:1075-1140) and PI-PI (1153-1258):
    zero resting length, the component along the fiber direction t (frozen)
    removed: E = w/2 |(I - t t') (B - A)|^2."""
    rng = np.random.default_rng(11 + dim)
    m, x0 = _meca(oracle, dim)
    if target == "exact":
        m.inter_sliding_link_ie(A1, A2, CA, B, W)
    else:
        m.inter_sliding_link_ii(A1, A2, CA, B1, B2, CB, W)
    M, bas = _recorded(oracle, m, dim)
    assert not bas.any()
    t = _pt(x0, dim, A2) - _pt(x0, dim, A1)
    t /= np.linalg.norm(t)
    P = np.eye(dim) - np.outer(t, t)

    def E(x):
        Bp = _pt(x, dim, B) if target == "exact" else _interp(x, dim, B1, B2, CB)
        v = P @ (Bp - _interp(x, dim, A1, A2, CA))
        return 0.5 * W * v @ v
    _check_quadratic(M, bas, E, x0, rng)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("target", ["exact", "interpolated"])
def test_side_sliding_links(oracle, dim, target):
    """interSideSlidingLink PI-PE (meca_inter.cc:1266-1499) and PI-PI
    (1508-1751): a side link whose force along the fiber is removed.
    2D: E = w/2 |(I - t t') (B - S)|^2, S = A + e R90 (P2 - P1);
    3D: only the arm direction b is kept: E = w/2 (b . (B - A - arm))^2."""
    rng = np.random.default_rng(21 + dim)
    m, x0 = _meca(oracle, dim)

    def tgt(x):
        return _pt(x, dim, B) if target == "exact" else _interp(x, dim, B1, B2, CB)

    d0 = _pt(x0, dim, A2) - _pt(x0, dim, A1)
    A0 = _interp(x0, dim, A1, A2, CA)
    length = 0.25
    s, off = _side_frame(dim, d0, tgt(x0) - A0, length)
    _assert_side_point(off, d0, tgt(x0) - A0, length)
    if target == "exact":
        m.inter_side_sliding_link_ie(A1, A2, CA, B, length, W)
    else:
        m.inter_side_sliding_link_ii(A1, A2, CA, B1, B2, CB, length, W)
    M, bas = _recorded(oracle, m, dim)
    if dim == 2:
        e = s * length / np.linalg.norm(d0)
        t = d0 / np.linalg.norm(d0)
        P = np.eye(2) - np.outer(t, t)

        def E(x):
            S = _interp(x, 2, A1, A2, CA) + e * _r90(_pt(x, 2, A2) - _pt(x, 2, A1))
            v = P @ (tgt(x) - S)
            return 0.5 * W * v @ v
    else:
        b = off / length

        def E(x):
            return 0.5 * W * (b @ (tgt(x) - _interp(x, 3, A1, A2, CA) - off)) ** 2
    _check_quadratic(M, bas, E, x0, rng)


@pytest.mark.parametrize("angle", [0.0, 0.7, -2.0])
def test_torque(oracle, angle):
    """interTorque2D (meca_inter.cc:1772-1869): E = w/2 d^2 with d the angle
    from AB to CD turned by -angle.  The recorded force is -grad E at x0 and
    the matrix its exact Hessian (w h h' plus w d times the derivative of the
    normals, meca_inter.cc:1826-1852).  No 3D version."""
    m3, _ = _meca(oracle, 3)
    with pytest.raises(ValueError):
        m3.inter_torque(A1, A2, CA, B1, B2, CB, 1.0, 0.0, W)
    m, x0 = _meca(oracle, 2)
    co, si = np.cos(angle), np.sin(angle)
    m.inter_torque(A1, A2, CA, B1, B2, CB, co, si, W)
    M, bas = _recorded(oracle, m, 2)

    def E(x):
        ab = _pt(x, 2, A2) - _pt(x, 2, A1)
        cd = _pt(x, 2, B2) - _pt(x, 2, B1)
        th = np.arctan2(cd[1], cd[0]) - angle - np.arctan2(ab[1], ab[0])
        d = np.angle(np.exp(1j * th))
        return 0.5 * W * d * d

    G = -_grad(E, x0, 1e-6)
    assert np.abs(M @ x0 + bas - G).max() < 1e-6 * W
    assert np.abs(M + _hess(E, x0)).max() < 1e-4 * W
    # the coefficients of the interpolation play no part
    m2, _ = _meca(oracle, 2)
    m2.inter_torque(A1, A2, 0.9, B1, B2, 0.1, co, si, W)
    M2, bas2 = _recorded(oracle, m2, 2)
    assert np.array_equal(M, M2) and np.array_equal(bas, bas2)


@pytest.mark.parametrize("dim", [2, 3])
def test_side_clamp(oracle, dim):
    """interSideClamp (meca_inter.cc:2032-2226): S = A + len N towards g, force
    w (g - S).  2D: E = w/2 |g - S|^2, S = A + e R90 (P2 - P1).  3D, segments of
    unit length: the arm n is orthogonal to the fiber and to g - A, S = A +
    e x (P2 - P1) with e = len n / L, and the component along n is removed,
    E = w/2 |(I - n n') (g - S)|^2."""
    rng = np.random.default_rng(31 + dim)
    m, x0 = _meca(oracle, dim, cut=1.0)
    g = _interp(x0, dim, A1, A2, CA) + np.array([0.2, 0.9, -0.3][:dim])
    d0 = _pt(x0, dim, A2) - _pt(x0, dim, A1)
    A0 = _interp(x0, dim, A1, A2, CA)
    length = 0.15
    s, off = _side_frame(dim, d0, g - A0, length)
    _assert_side_point(off, d0, g - A0, length)
    m.inter_side_clamp(A1, A2, CA, g, length, W)
    M, bas = _recorded(oracle, m, dim)
    if dim == 2:
        e = s * length / np.linalg.norm(d0)

        def E(x):
            S = _interp(x, 2, A1, A2, CA) + e * _r90(_pt(x, 2, A2) - _pt(x, 2, A1))
            return 0.5 * W * np.sum((g - S) ** 2)
    else:
        n = np.cross(A0 - g, d0)
        n /= np.linalg.norm(n)
        ev = length * n / np.linalg.norm(d0)
        P = np.eye(3) - np.outer(n, n)

        def E(x):
            S = _interp(x, 3, A1, A2, CA) + np.cross(ev, _pt(x, 3, A2) - _pt(x, 3, A1))
            v = P @ (g - S)
            return 0.5 * W * v @ v
        assert np.abs(np.cross(ev, d0) - off).max() < 1e-12
    _check_quadratic(M, bas, E, x0, rng)


def test_side_clamp_3d_scales_the_arm_term_by_the_segment_length(oracle):
    """interSideClamp3D (meca_inter.cc:2153-2160) removes the arm component with
    TP = (aa, bb) (x) arm / (|arm| L): for a segment of length L the matrix is
    -w T' (I - n n' / L^2) T while vBAS is w T' (I - n n') g, T the map from the
    two points to S.  Checked at L = 0.5, with T from S(x) itself."""
    m, x0 = _meca(oracle, 3, cut=0.5)
    g = _interp(x0, 3, A1, A2, CA) + np.array([0.2, 0.9, -0.3])
    d0 = _pt(x0, 3, A2) - _pt(x0, 3, A1)
    L = np.linalg.norm(d0)
    A0 = _interp(x0, 3, A1, A2, CA)
    length = 0.15
    m.inter_side_clamp(A1, A2, CA, g, length, W)
    M, bas = _recorded(oracle, m, 3)
    n = np.cross(A0 - g, d0)
    n /= np.linalg.norm(n)
    ev = length * n / L

    def S(x):
        return _interp(x, 3, A1, A2, CA) + np.cross(ev, _pt(x, 3, A2) - _pt(x, 3, A1))
    T = np.stack([S(e) - S(np.zeros(x0.size)) for e in np.eye(x0.size)], axis=1)
    assert np.abs(M + W * T.T @ (np.eye(3) - np.outer(n, n) / L ** 2) @ T).max() < 1e-12 * W / L ** 2
    assert np.abs(bas - W * T.T @ (np.eye(3) - np.outer(n, n)) @ g).max() < 1e-12 * W


@pytest.mark.parametrize("dim", [2, 3])
def test_sliding_clamp_and_plane(oracle, dim):
    """interSlidingClamp (meca_inter.cc:2241-2283): E = w/2 |(I - t t')(g - A)|^2;
    interPlane(PointInterpolated) (2340-2378): E = w/2 (n . (g - A))^2."""
    rng = np.random.default_rng(41 + dim)
    g = np.array([0.4, -0.2, 0.7][:dim])
    nv = rng.standard_normal(dim)
    nv /= np.linalg.norm(nv)
    m, x0 = _meca(oracle, dim)
    m.inter_sliding_clamp(A1, A2, CA, g, W)
    M, bas = _recorded(oracle, m, dim)
    t = _pt(x0, dim, A2) - _pt(x0, dim, A1)
    t /= np.linalg.norm(t)
    P = np.eye(dim) - np.outer(t, t)

    def E1(x):
        v = P @ (g - _interp(x, dim, A1, A2, CA))
        return 0.5 * W * v @ v
    _check_quadratic(M, bas, E1, x0, rng)
    m, x0 = _meca(oracle, dim)
    m.inter_plane_i(A1, A2, CA, nv, g, W)
    M, bas = _recorded(oracle, m, dim)

    def E2(x):
        return 0.5 * W * (nv @ (g - _interp(x, dim, A1, A2, CA))) ** 2
    _check_quadratic(M, bas, E2, x0, rng)


@pytest.mark.parametrize("dim", [2, 3])
def test_overlapping_points_record_nothing(oracle, dim):
    """PointInterpolated::overlapping (point_interpolated.cc:21-33): a link
    between points that share a model point of one Mecable is skipped."""
    m, x0 = _meca(oracle, dim)
    m.inter_long_link_ie(A1, A2, CA, A2, 0.3, W)
    m.inter_side_link_ii(A1, A2, CA, A2, A2 + 1, CB, 0.3, W)
    m.inter_sliding_link_ie(A1, A2, CA, A1, W)
    m.inter_side_sliding_link_ii(A1, A2, CA, A2, A2 + 1, CB, 0.3, W)
    m.inter_long_link_ii(A1, A2, CA, A1 - 1, A1, CB, 0.3, W)
    M, bas = _recorded(oracle, m, dim)
    assert not M.any() and not bas.any()
