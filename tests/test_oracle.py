"""Pins the CPU oracle (oracle/meca_oracle.c) before it is trusted as the checker.

The reference holds no golden vectors for this path and its own build needs an
external LAPACK and its build system, so every function of the oracle is pinned
here against an independent route: scipy's LAPACK for the restated routines,
closed forms and invariants (projection idempotence, bending null space,
rank-1 link blocks, force-law linearisation), dense brute force for the
operators, and dense solves for the Krylov solver.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as LP

from conftest import oracle_from_scene, rel_err


# ---------------------------------------------------------------- LAPACK ---

@pytest.mark.parametrize("n", [1, 2, 3, 7, 31, 93])
def test_dgetrf_dgetri_match_scipy(oracle, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + (n > 2) * np.diag(rng.standard_normal(n))
    lu, ipiv, info = oracle.dgetrf(A)
    lu2, piv2, info2 = LP.dgetrf(A)
    assert info == info2 == 0
    assert np.array_equal(ipiv - 1, piv2)
    assert np.abs(lu - lu2).max() <= 1e-12 * np.abs(lu2).max()
    inv, info = oracle.dgetri(lu, ipiv)
    assert info == 0
    assert rel_err(inv, np.linalg.inv(A)) < 1e-9


def test_dgetrf_zero_pivot_reported(oracle):
    A = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [0.0, 0.0, 0.0]])
    _, _, info = oracle.dgetrf(A)
    _, _, info2 = LP.dgetrf(A)
    assert info == info2 > 0


def test_pttrf_ptts2_match_scipy(oracle):
    rng = np.random.default_rng(1)
    for n in (2, 5, 30):
        d = 2 + rng.random(n)
        e = -0.9 * rng.random(n - 1)
        dd, ee, info = oracle.dpttrf(d, e)
        d2, e2, info2 = LP.dpttrf(d, e)
        assert info == info2 == 0
        assert np.allclose(dd, d2, rtol=1e-15, atol=0) and np.allclose(ee, e2, rtol=1e-15, atol=0)
        b = rng.standard_normal(n)
        x = oracle.dptts2(dd, ee, b)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert np.abs(T @ x - b).max() < 1e-12


def test_pttrf_rejects_indefinite(oracle):
    _, _, info = oracle.dpttrf([1.0, 1.0], [2.0])
    assert info == 2


def test_gtsv_matches_scipy_with_pivoting(oracle):
    rng = np.random.default_rng(2)
    for n in (2, 3, 8, 40):
        dl = rng.standard_normal(n - 1) * 3.0       # large subdiagonal forces row interchanges
        d = rng.standard_normal(n)
        du = rng.standard_normal(n - 1)
        b = rng.standard_normal(n)
        x, info = oracle.dgtsv(dl, d, du, b)
        r = LP.dgtsv(dl, d, du, b)
        assert info == r[4] == 0
        assert np.abs(x - r[3]).max() <= 1e-12 * max(1, np.abs(r[3]).max())


# ------------------------------------------------------------ fiber model ---

def test_best_nb_points_is_the_argmin(oracle):
    for ratio in np.concatenate([np.linspace(0.05, 40, 4001), [1.0, 2.0, 22.0, 2.5, 4.5]]):
        N = np.arange(1, 200)
        best = N[np.argmin(np.abs(ratio / N - 1.0))] + 1
        assert oracle.best_nb_points(ratio) == best


def _one_fiber(oracle, dim, pts, cut, rigidity=20.0, visc=0.02):
    m = oracle.OracleMeca(dim, dict(time_step=0.005, kT=0.0042, viscosity=visc))
    m.add_fiber(len(pts) // dim, cut, rigidity)
    m.prepare(pts)
    return m


def _bent_fiber(dim, np_, cut, seed=0):
    rng = np.random.default_rng(seed)
    d = 0.4 * rng.standard_normal((np_ - 1, dim)) + np.eye(dim)[0]
    d = cut * d / np.linalg.norm(d, axis=1, keepdims=True)
    P = np.vstack([np.zeros(dim), np.cumsum(d, axis=0)])
    return P.ravel()


@pytest.mark.parametrize("L", [0.01, 0.3, 2.0, 8.5, 10.0, 15.0, 40.0])
def test_mobility_tirado_de_la_torre(oracle, L):
    """rfMobility = points / drag, drag = pi*eta*max(3L/(ln(p)+0.312), 6 r0) with
    p = min(L, r1)/(2 r0) (Tirado & de la Torre 1979; sphere floor); the
    cut-off r1 = 10 um applies to the logarithm only."""
    eta, r0, r1 = 0.02, 0.0125, 10.0
    npts = 5
    cut = L / (npts - 1)
    m = _one_fiber(oracle, 3, _bent_fiber(3, npts, cut), cut, visc=eta)
    p = min(L, r1) / (2 * r0)
    cyl = 3 * np.pi * eta * L / (np.log(p) + 0.312)
    drag = max(cyl, 6 * np.pi * eta * r0)
    assert m.mobility(0) == pytest.approx(npts / drag, rel=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_projection_is_the_orthogonal_projector(oracle, dim):
    np_, cut = 9, 0.5
    pts = _bent_fiber(dim, np_, cut, seed=dim)
    m = _one_fiber(oracle, dim, pts, cut)
    n = dim * np_
    P = np.stack([m.project(0, e) for e in np.eye(n)], axis=1)
    X = pts.reshape(np_, dim)
    J = np.zeros((np_ - 1, n))
    for s in range(np_ - 1):
        ds = (X[s + 1] - X[s]) / cut
        J[s, dim * s:dim * s + dim] = -ds
        J[s, dim * (s + 1):dim * (s + 1) + dim] = ds
    Pref = np.eye(n) - J.T @ np.linalg.solve(J @ J.T, J)
    assert np.abs(P - Pref).max() < 1e-12
    assert np.abs(P @ P - P).max() < 1e-12
    assert np.abs(P - P.T).max() < 1e-12
    assert np.abs(P @ J.T).max() < 1e-12
    d, e = m.jjt(0)
    JJ = J @ J.T
    assert np.allclose(JJ.diagonal(), np.concatenate([[d[0]], d[1:] + e * e * d[:-1]]), rtol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("np_", [2, 3, 4, 5, 12])
def test_bending_matrix(oracle, dim, np_):
    """R = -k (D'D) x I_dim, D the second difference; annihilates affine modes."""
    cut, k = 0.5, 20.0
    m = _one_fiber(oracle, dim, _bent_fiber(dim, np_, cut), cut, rigidity=k)
    n = dim * np_
    R = np.stack([m.add_rigidity(0, e) for e in np.eye(n)], axis=1)
    if np_ < 3:
        assert not R.any()
        return
    Dm = np.zeros((np_ - 2, np_))
    for j in range(np_ - 2):
        Dm[j, j:j + 3] = [1, -2, 1]
    Rref = -(k / cut ** 3) * np.kron(Dm.T @ Dm, np.eye(dim))
    assert np.abs(R - Rref).max() < 1e-9
    affine = (np.arange(np_)[:, None] * np.linspace(0.3, 0.7, dim)[None, :] + 1.0).ravel()
    assert np.abs(R @ affine).max() < 1e-9


def test_block_of_isolated_fiber_uses_mat_up(oracle):
    """getBlock (via addRigidityMatUp) equals the dense operator's block."""
    dim, np_, cut = 3, 7, 0.5
    m = _one_fiber(oracle, dim, _bent_fiber(dim, np_, cut), cut)
    m.setup(np.random.default_rng(0).standard_normal(dim * np_))
    n = dim * np_
    A = np.stack([m.multiply(e) for e in np.eye(n)], axis=1)
    assert np.abs(m.get_block(0) - A).max() < 1e-12


# ------------------------------------------------------------- assembly ---

@pytest.mark.parametrize("dim", [2, 3])
def test_links_are_rank_one_blocks(oracle, scenes, dim):
    """interLink(PI,PI) adds -w c c' with c = [1-a, a, -(1-b), -b]."""
    sc = scenes.mini(dim=dim, seed=31, nfib=6, nlinks=15)
    sc.confine_radius = 0.0
    m = oracle_from_scene(oracle, sc)
    B = m.matrix_dense(0)
    a1, a2, ca = sc.interpolate(sc.link_fiber_a, sc.link_abs_a)
    b1, b2, cb = sc.interpolate(sc.link_fiber_b, sc.link_abs_b)
    Bref = np.zeros_like(B)
    for k in range(sc.nlinks):
        c = np.zeros(sc.npts)
        c[a1[k]] += 1 - ca[k]; c[a2[k]] += ca[k]; c[b1[k]] -= 1 - cb[k]; c[b2[k]] -= cb[k]
        Bref -= sc.link_stiffness[k] * np.outer(c, c)
    assert np.abs(B - Bref).max() < 1e-11
    assert np.abs(B.sum(axis=1)).max() < 1e-10


def test_sparse_layout_is_sorted_unique_lower(oracle, scenes):
    sc = scenes.mini(dim=3, seed=32, nfib=6, nlinks=15)
    m = oracle_from_scene(oracle, sc)
    ija, sa = m.matrix_nr(0)
    n = sc.npts
    assert ija[0] == n + 1
    touched = set()
    a1, a2, _ = sc.interpolate(sc.link_fiber_a, sc.link_abs_a)
    b1, b2, _ = sc.interpolate(sc.link_fiber_b, sc.link_abs_b)
    for k in range(sc.nlinks):
        idx = [a1[k], a2[k], b1[k], b2[k]]
        for i in idx:
            for j in idx:
                if i > j:
                    touched.add((int(i), int(j)))
    got = set()
    for j in range(n):
        rows = ija[ija[j]:ija[j + 1]]
        assert np.all(np.diff(rows.astype(np.int64)) > 0) and np.all(rows > j)
        got.update((int(r), j) for r in rows)
    assert got == touched


@pytest.mark.parametrize("dim", [2, 3])
def test_long_clamp_linearises_the_force_law(oracle, dim):
    """interLongClamp outside the radius: F(x) = w (len/|x| - 1) x; the matrix
    block and vBAS reproduce F at x0 and its Jacobian (finite differences)."""
    rng = np.random.default_rng(dim)
    np_, cut, w, R = 3, 0.5, 200.0, 1.0
    pts = _bent_fiber(dim, np_, cut) + 2.0
    m = _one_fiber(oracle, dim, pts, cut)
    x0 = pts[:dim].copy()
    m.inter_long_clamp(0, np.zeros(dim), R, w)
    C = m.matrix_dense(1)[:dim, :dim]
    m.setup(None)
    bas = m.vector(oracle.BAS)[:dim]

    def F(x):
        return w * (R / np.linalg.norm(x) - 1.0) * x

    assert np.abs(C @ x0 + bas - F(x0)).max() < 1e-10 * w
    h = 1e-6
    Jfd = np.stack([(F(x0 + h * e) - F(x0 - h * e)) / (2 * h) for e in np.eye(dim)], axis=1)
    assert np.abs(C - Jfd).max() < 1e-5 * w


def _two_fiber_meca(oracle, dim):
    pts = np.concatenate([_bent_fiber(dim, 5, 0.5, seed=1), _bent_fiber(dim, 4, 0.5, seed=2) + 1.0])
    m = oracle.OracleMeca(dim, dict(time_step=0.005, kT=0.0042, viscosity=0.02))
    m.add_fiber(5, 0.5, 20.0)
    m.add_fiber(4, 0.5, 20.0)
    m.prepare(pts)
    return m, pts


def _linear_force(oracle, m, dim, x):
    """vBAS + (mB (x) I + mC) x: the force law an emitter recorded."""
    B = m.matrix_dense(0)
    C = m.matrix_dense(1)
    m.setup(None)
    bas = m.vector(oracle.BAS)
    return np.kron(B, np.eye(dim)) @ x + C @ x + bas


def _pt(x, dim, i):
    return x[dim * i:dim * i + dim]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", ["link_ee", "link_ie", "clamp_e", "clamp_i", "plane_e", "pure_force"])
def test_linear_emitters_follow_their_force_law(oracle, dim, kind):
    """Each Hookean emitter (meca_inter.cc: interLink PE-PE 196-230, PI-PE
    240-275, interClamp PE 1883-1907 and PI, interPlane PE 2301-2327,
    addPureForce) records a force that is exactly linear in the positions:
    vBAS + (mB (x) I + mC) x equals the spring force at any x."""
    rng = np.random.default_rng(7 + dim)
    m, pts = _two_fiber_meca(oracle, dim)
    n = pts.size
    w = 37.5
    g = rng.standard_normal(dim)
    nrm = rng.standard_normal(dim)
    nrm /= np.linalg.norm(nrm)
    a1, a2, ca, b = 1, 2, 0.3, 6
    if kind == "link_ee":
        m.inter_link_ee(a1, b, w)
    elif kind == "link_ie":
        m.inter_link_ie(a1, a2, ca, b, w)
    elif kind == "clamp_e":
        m.inter_clamp_e(b, g, w)
    elif kind == "clamp_i":
        m.inter_clamp_i(a1, a2, ca, g, w)
    elif kind == "plane_e":
        m.inter_plane_e(b, nrm, g, w)
    else:
        m.add_pure_force(b, g)
    for trial in range(3):
        x = pts + 0.3 * rng.standard_normal(n)
        F = np.zeros(n)
        if kind == "link_ee":
            f = w * (_pt(x, dim, b) - _pt(x, dim, a1))
            F[dim * a1:dim * a1 + dim] += f
            F[dim * b:dim * b + dim] -= f
        elif kind == "link_ie":
            pa = (1 - ca) * _pt(x, dim, a1) + ca * _pt(x, dim, a2)
            f = w * (_pt(x, dim, b) - pa)
            F[dim * a1:dim * a1 + dim] += (1 - ca) * f
            F[dim * a2:dim * a2 + dim] += ca * f
            F[dim * b:dim * b + dim] -= f
        elif kind == "clamp_e":
            F[dim * b:dim * b + dim] += w * (g - _pt(x, dim, b))
        elif kind == "clamp_i":
            f = w * (g - ((1 - ca) * _pt(x, dim, a1) + ca * _pt(x, dim, a2)))
            F[dim * a1:dim * a1 + dim] += (1 - ca) * f
            F[dim * a2:dim * a2 + dim] += ca * f
        elif kind == "plane_e":
            F[dim * b:dim * b + dim] += w * nrm * (nrm @ (g - _pt(x, dim, b)))
        else:
            F[dim * b:dim * b + dim] += g
        assert np.abs(_linear_force(oracle, m, dim, x) - F).max() < 1e-11 * max(w, 1.0)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("stretch", [1.6, 0.6])
def test_long_link_linearisation(oracle, dim, stretch):
    """interLongLink PE-PE (meca_inter.cc:334-405): F = w (|ab| - len) ab/|ab|.
    The recorded terms give the exact force at the current positions; when the
    link is stretched the matrix is the exact Jacobian (finite differences),
    when compressed it is the rank-1 negative semidefinite -w u u' block that
    removes the negative eigenvalues (meca_inter.cc:372-374)."""
    rng = np.random.default_rng(3)
    m, pts = _two_fiber_meca(oracle, dim)
    a, b, w = 1, 6, 50.0
    d0 = np.linalg.norm(_pt(pts, dim, b) - _pt(pts, dim, a))
    length = d0 / stretch
    m.inter_long_link_ee(a, b, length, w)

    def force(x):
        ab = _pt(x, dim, b) - _pt(x, dim, a)
        nab = np.linalg.norm(ab)
        f = w * (nab - length) * ab / nab
        F = np.zeros(x.size)
        F[dim * a:dim * a + dim] += f
        F[dim * b:dim * b + dim] -= f
        return F

    assert np.abs(_linear_force(oracle, m, dim, pts) - force(pts)).max() < 1e-10 * w
    C = m.matrix_dense(1)
    ia = slice(dim * a, dim * a + dim)
    if stretch > 1:
        h = 1e-6
        J = np.stack([(force(pts + h * e) - force(pts - h * e)) / (2 * h) for e in np.eye(pts.size)], axis=1)
        assert np.abs(C - J).max() < 1e-5 * w
    else:
        u = _pt(pts, dim, b) - _pt(pts, dim, a)
        u /= np.linalg.norm(u)
        assert np.abs(C[ia, ia] + w * np.outer(u, u)).max() < 1e-10 * w
        assert np.linalg.eigvalsh(C[ia, ia]).max() < 1e-10 * w


@pytest.mark.parametrize("dim", [2, 3])
def test_operator_equals_dense_construction(oracle, scenes, dim):
    """Meca::multiply == I - dt*mob*P (B x I + C + R + P') built from the pieces."""
    sc = scenes.mini(dim=dim, seed=33, nfib=5, nlinks=12)
    m = oracle_from_scene(oracle, sc)
    m.setup(sc.gauss(0))
    n = sc.size
    A = np.stack([m.multiply(e) for e in np.eye(n)], axis=1)
    K = np.kron(m.matrix_dense(0), np.eye(dim)) + m.matrix_dense(1)
    Pm = np.zeros((n, n))
    Rm = np.zeros((n, n))
    Dm = np.zeros((n, n))
    for f in range(sc.nfib):
        o, bs = dim * sc.mat_index[f], dim * sc.nb_points[f]
        E = np.eye(bs)
        Pm[o:o + bs, o:o + bs] = np.stack([m.project(f, e, 1.0) for e in E], axis=1)
        Rm[o:o + bs, o:o + bs] = np.stack([m.add_rigidity(f, e) for e in E], axis=1)
        Dm[o:o + bs, o:o + bs] = np.stack([m.add_projection_diff(f, e) for e in E], axis=1)
    mob = np.concatenate([np.full(dim * sc.nb_points[f], m.mobility(f)) for f in range(sc.nfib)])
    Aref = np.eye(n) - sc.time_step * mob[:, None] * (Pm @ (K + Rm + Dm))
    assert np.abs(A - Aref).max() < 1e-10 * np.abs(A).max()
    # P' is symmetric negative semidefinite and blind to translations
    assert np.abs(Dm - Dm.T).max() < 1e-12
    assert np.linalg.eigvalsh(Dm).max() < 1e-9
    # getBlock equals the operator's diagonal block; its inverse inverts it
    m.compute_preconditioner()
    for f in range(sc.nfib):
        o, bs = dim * sc.mat_index[f], dim * sc.nb_points[f]
        blk = m.get_block(f)
        assert np.abs(blk - A[o:o + bs, o:o + bs]).max() < 1e-10 * np.abs(blk).max()
        assert np.abs(m.get_inverse_block(f) @ blk - np.eye(bs)).max() < 1e-10


def test_brownian_terms_and_noise_level(oracle, scenes):
    sc = scenes.mini(dim=3, seed=34, nfib=5, nlinks=10)
    m = oracle_from_scene(oracle, sc)
    g = sc.gauss(0)
    m.setup(g)
    F0 = m.forces_linear(m.vector(oracle.PTS), True)
    dF = m.vector(oracle.FOR) - F0
    sc_kT = sc.kT / sc.time_step
    noise = np.inf
    for f in range(sc.nfib):
        o, bs = 3 * sc.mat_index[f], 3 * sc.nb_points[f]
        b = np.sqrt(2 * sc_kT / m.mobility(f))
        assert np.allclose(dF[o:o + bs], b * g[o:o + bs], rtol=1e-9, atol=1e-9)
        noise = min(noise, m.mobility(f) * b)
    rep = m.iterate(0.05)
    assert rep["noise_level"] == pytest.approx(noise * sc.time_step, rel=1e-14)
    assert rep["threshold"] == pytest.approx(0.05 * noise * sc.time_step, rel=1e-14)


# ---------------------------------------------------------------- solver ---

@pytest.mark.parametrize("precondition", [0, 1])
def test_bicgstab_solves_the_dense_system(oracle, scenes, precondition):
    sc = scenes.mini(dim=3, seed=35, nfib=6, nlinks=15)
    m = oracle_from_scene(oracle, sc)
    m.setup(sc.gauss(0))
    n = sc.size
    A = np.stack([m.multiply(e) for e in np.eye(n)], axis=1)
    b = m.vector(oracle.RHS)
    x = np.linalg.solve(A, b)
    rep = m.iterate(1e-12, tol_mode=1, precondition=precondition)
    assert rep["converged"]
    assert rel_err(m.vector(oracle.SOL), x) < 1e-9
    # the converged solution does not depend on the initial guess
    rep2 = m.iterate(1e-12, tol_mode=1, precondition=precondition, guess=np.ones(n))
    assert rep2["converged"] and rel_err(m.vector(oracle.SOL), x) < 1e-9


def test_config1_needs_one_iteration(oracle, scenes):
    """Uncoupled fibers: the block preconditioner is the exact inverse, BCGSP
    stops after one iteration (as measured with the compiled reference)."""
    sc = scenes.config1(seed=1)
    m = oracle_from_scene(oracle, sc)
    rep = m.solve(0.05, sc.gauss(0))
    assert rep["iterations"] == 1 and rep["converged"]


def test_rescue_chain_and_failure(oracle, scenes):
    sc = scenes.mini(dim=3, seed=36, nfib=6, nlinks=15)
    m = oracle_from_scene(oracle, sc)
    m.setup(sc.gauss(0))
    rep = m.iterate(1e-14, tol_mode=1, max_iter=1)
    assert rep["failed"] and rep["fallback"] == 3


# --------------------------------------------------------------- reshape ---

@pytest.mark.parametrize("dim", [2, 3])
def test_reshape_restores_segment_lengths(oracle, dim):
    rng = np.random.default_rng(7)
    cut = 0.5
    for np_ in (2, 3, 8, 31):
        pts = _bent_fiber(dim, np_, cut, seed=np_) + 0.01 * rng.standard_normal(dim * np_)
        out, failed = oracle.reshape(dim, pts, cut)
        assert not failed, "Newton (reshape_it) converges for a 2% perturbation"
        X = out.reshape(np_, dim)
        seg = np.linalg.norm(np.diff(X, axis=0), axis=1)
        assert np.abs(seg - cut).max() < 1e-6
        # zero-momentum correction: the centroid does not move (to Newton accuracy)
        assert np.abs(X.mean(0) - pts.reshape(np_, dim).mean(0)).max() < 1e-3


def test_reshape_sure_fallback_keeps_centroid(oracle):
    """A fiber far from its constraint makes reshape_it fail; reshape_sure then
    restores the lengths segment by segment and moves the centroid back."""
    dim, np_, cut = 2, 31, 0.5
    rng = np.random.default_rng(3)
    pts = np.cumsum(rng.uniform(0.05, 0.9, (np_, dim)), axis=0).ravel()
    out, failed = oracle.reshape(dim, pts, cut)
    assert failed
    X = out.reshape(np_, dim)
    assert np.abs(np.linalg.norm(np.diff(X, axis=0), axis=1) - cut).max() < 1e-12
    assert np.abs(X.mean(0) - pts.reshape(np_, dim).mean(0)).max() < 1e-12


# ----------------------------------------- multipliers, P', reshape form ---
# These pin the three oracle functions that the operator construction above
# only checks for consistency with the oracle's own pieces: the Lagrange
# multipliers (rfLag, rigid_fiber_project.cc:137-161), the P' coefficients
# built from them (makeProjectionDiff, rigid_fiber_project.cc:168-219) and the
# Newton-on-multipliers form of FiberNaked::reshape_it (fiber_naked.cc:221-354).

def _straight_fiber(dim, np_, cut, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(dim)
    u /= np.linalg.norm(u)
    x0 = rng.standard_normal(dim)
    return (x0[None, :] + cut * np.arange(np_)[:, None] * u[None, :]).ravel(), u


def _forced_fiber(oracle, dim, pts, cut, forces):
    """One fiber, no Brownian term (kT = 0), external forces {point: vector}
    placed in vBAS; rigidity 0 so that vFOR holds exactly these forces."""
    m = oracle.OracleMeca(dim, dict(time_step=0.005, kT=0.0, viscosity=0.02))
    m.add_fiber(len(pts) // dim, cut, 0.0)
    m.prepare(pts)
    for p, f in forces.items():
        m.add_pure_force(p, np.asarray(f, dtype=np.float64))
    m.setup(None)
    return m


def _chain_laplacian(np_, coef, dim):
    """Matrix of Y_s += c_s (X_{s+1} - X_s), Y_{s+1} -= c_s (X_{s+1} - X_s)."""
    L = np.zeros((np_, np_))
    for s, c in enumerate(coef):
        L[s, s] -= c
        L[s, s + 1] += c
        L[s + 1, s] += c
        L[s + 1, s + 1] -= c
    return np.kron(L, np.eye(dim))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("np_", [2, 3, 9, 31])
def test_multipliers_and_projection_diff_on_a_pulled_straight_fiber(oracle, dim, np_):
    """A straight fiber pulled apart by +-F at its two ends carries the tension F
    in every segment, so every multiplier is F and every P' coefficient F/cut
    (P' = -(F/cut) * chain Laplacian). Pushed together, the multipliers are -F
    and the compressive multipliers are dropped: P' = 0."""
    cut, F = 0.5, 3.25
    pts, u = _straight_fiber(dim, np_, cut, seed=np_ + dim)
    n = dim * np_
    for sign in (+1.0, -1.0):
        m = _forced_fiber(oracle, dim, pts, cut, {0: -sign * F * u, np_ - 1: sign * F * u})
        lag = m.tensions()
        assert np.abs(lag - sign * F).max() < 1e-12 * F
        D = np.stack([m.add_projection_diff(0, e) for e in np.eye(n)], axis=1)
        c = F / cut if sign > 0 else 0.0
        assert np.abs(D - _chain_laplacian(np_, [c] * (np_ - 1), dim)).max() < 1e-12 * F / cut


@pytest.mark.parametrize("dim", [2, 3])
def test_projection_diff_keeps_tension_and_drops_compression(oracle, dim):
    """Held at its middle point m, with the minus end pushed inward and the
    plus end pulled outward (forces F u, -2F u, F u at points 0, m, last): the
    segments left of m are compressed (multiplier -F), those right of m are
    stretched (+F). P' keeps only the stretched half, at F/cut."""
    np_, cut, F = 11, 0.5, 1.5
    mid = 5
    pts, u = _straight_fiber(dim, np_, cut, seed=7)
    m = _forced_fiber(oracle, dim, pts, cut, {0: F * u, mid: -2 * F * u, np_ - 1: F * u})
    lag = m.tensions()
    expect = np.where(np.arange(np_ - 1) < mid, -F, F)
    assert np.abs(lag - expect).max() < 1e-12 * F
    D = np.stack([m.add_projection_diff(0, e) for e in np.eye(dim * np_)], axis=1)
    coef = np.where(np.arange(np_ - 1) < mid, 0.0, F / cut)
    assert np.abs(D - _chain_laplacian(np_, coef, dim)).max() < 1e-12 * F / cut


def _constraint_jacobian(X, cut):
    """J with rows d_s = (x_{s+1} - x_s)/cut acting on (x_{s+1} - x_s)."""
    np_, dim = X.shape
    J = np.zeros((np_ - 1, dim * np_))
    for s in range(np_ - 1):
        ds = (X[s + 1] - X[s]) / cut
        J[s, dim * s:dim * s + dim] = -ds
        J[s, dim * (s + 1):dim * (s + 1) + dim] = ds
    return J


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("np_", [3, 8, 23])
def test_tensions_equal_dense_multipliers(oracle, dim, np_):
    """rfLag = (J J^T)^-1 J F, with J from the fiber's own points and the
    normal equations solved densely, for random forces on a bent fiber; both
    from the right-hand side (setSpeedsFromForces, rhs = true) and from
    Meca::computeForces (computeTensions)."""
    cut = 0.5
    rng = np.random.default_rng(100 * dim + np_)
    pts = _bent_fiber(dim, np_, cut, seed=np_) * 1.07      # segments 7% longer than cut
    forces = {p: rng.standard_normal(dim) for p in range(np_)}
    m = _forced_fiber(oracle, dim, pts, cut, forces)
    Fv = np.concatenate([forces[p] for p in range(np_)])
    J = _constraint_jacobian(pts.reshape(np_, dim), cut)
    lam = np.linalg.solve(J @ J.T, J @ Fv)
    assert np.abs(m.tensions() - lam).max() < 1e-12 * np.abs(lam).max()
    m.compute_forces()
    assert np.abs(m.tensions() - lam).max() < 1e-12 * np.abs(lam).max()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("np_", [3, 8, 31])
def test_reshape_it_moves_along_the_constraint_gradients(oracle, dim, np_):
    """reshape_it solves for multipliers sigma: the new points are
    src + J(src)^T sigma with J from the segments at src, chosen so that every
    segment has length cut. So (1) the displacement lies in the range of
    J(src)^T (which also keeps the centroid exactly), and (2) it is the root of
    the length equations found here independently by scipy from sigma = 0."""
    from scipy.optimize import fsolve
    cut = 0.5
    rng = np.random.default_rng(50 + np_)
    src = _bent_fiber(dim, np_, cut, seed=np_ + 3) + 0.02 * rng.standard_normal(dim * np_)
    out, failed = oracle.reshape(dim, src, cut)
    assert not failed
    X = src.reshape(np_, dim)
    Jt = (_constraint_jacobian(X, cut) * cut).T          # columns x_{s+1} - x_s placed as in J
    delta = out - src
    sig, *_ = np.linalg.lstsq(Jt, delta, rcond=None)
    assert np.abs(Jt @ sig - delta).max() < 1e-12 * np.abs(delta).max()

    def lengths(s):
        Y = (src + Jt @ s).reshape(np_, dim)
        return (np.diff(Y, axis=0) ** 2).sum(1) - cut * cut

    s_ref = fsolve(lengths, np.zeros(np_ - 1))
    assert np.abs(lengths(s_ref)).max() < 1e-10
    assert np.abs((src + Jt @ s_ref) - out).max() < 1e-6
    # a reshape that is not of this form fails check (1): reshape_sure, the
    # fallback taken when Newton fails (fiber_naked.cc:367-401)
    far = np.cumsum(rng.uniform(0.05, 0.9, (31, 2)), axis=0).ravel()
    out2, failed2 = oracle.reshape(2, far, cut)
    assert failed2
    Jt2 = (_constraint_jacobian(far.reshape(31, 2), cut) * cut).T
    d2 = out2 - far
    sig2, *_ = np.linalg.lstsq(Jt2, d2, rcond=None)
    assert np.abs(Jt2 @ sig2 - d2).max() > 1e-3 * np.abs(d2).max()


@pytest.mark.parametrize("L", [0.3, 2.0, 8.5, 15.0])
@pytest.mark.parametrize("h", [0.005, 0.05, 1.0])
def test_mobility_surface_model(oracle, L, h):
    """Near a surface (fiber:surface_effect), drag = 2 pi eta L / acosh(1 + h/r)
    (Hunt et al. 1994, the formula of Fiber::dragCoefficientSurface): for
    x = 1 + h/r >> 1, acosh(x) = ln(2x) to O(x^-2); a per-fiber radius
    overrides the Meca's parameters."""
    eta, r = 0.02, 0.01
    npts = 6
    cut = L / (npts - 1)
    m = oracle.OracleMeca(3, dict(time_step=0.005, kT=0.0042, viscosity=eta))
    m.add_fiber(npts, cut, 20.0, hydro_radius0=r, surface_height=h)
    m.add_fiber(npts, cut, 20.0, hydro_radius0=r)
    pts = np.concatenate([_bent_fiber(3, npts, cut), _bent_fiber(3, npts, cut, seed=1) + 3])
    m.prepare(pts)
    x = 1 + h / r
    drag = 2 * np.pi * eta * L / np.log(x + np.sqrt(x * x - 1))
    assert m.mobility(0) == pytest.approx(npts / drag, rel=1e-13)
    if x > 50:
        assert m.mobility(0) == pytest.approx(npts * np.log(2 * x) / (2 * np.pi * eta * L), rel=1e-4)
    # volume model with the fiber's own radius (r = 0.01 instead of 0.0125)
    p = min(L, 10.0) / (2 * r)
    vol = np.pi * eta * max(3 * L / (np.log(p) + 0.312), 6 * r)
    assert m.mobility(1) == pytest.approx(npts / vol, rel=1e-13)


def test_couple_list_assembly_equals_one_link_per_call(oracle, scenes):
    """orc_inter_couples (one C call for the couple list, used by the CPU
    baseline) records exactly the interLink calls of the per-couple loop, in
    the same order: bit-identical mB; couples with a free hand are skipped."""
    sc = scenes.mini(dim=3, seed=61, nfib=10, nlinks=40)
    state = np.full(sc.nlinks, 3, np.uint8)
    state[::5] = 1
    m1 = oracle.OracleMeca(3, sc.params())
    m2 = oracle.OracleMeca(3, sc.params())
    for m in (m1, m2):
        oracle.register_scene(m, sc)
        m.prepare(sc.points)
    m1.inter_couples(sc.link_fiber_a, sc.link_abs_a, sc.link_fiber_b, sc.link_abs_b, sc.link_stiffness, state)
    a1, a2, ca = sc.interpolate(sc.link_fiber_a, sc.link_abs_a)
    b1, b2, cb = sc.interpolate(sc.link_fiber_b, sc.link_abs_b)
    for k in range(sc.nlinks):
        if state[k] == 3:
            m2.inter_link_ii(int(a1[k]), int(a2[k]), float(ca[k]), int(b1[k]), int(b2[k]), float(cb[k]),
                             float(sc.link_stiffness[k]))
    assert np.array_equal(m1.matrix_dense(0), m2.matrix_dense(0))
    assert np.any(m1.matrix_dense(0))
