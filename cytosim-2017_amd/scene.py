"""Synthetic fiber scenes for the implicit mechanics step.

A scene is the input one `Simul::solve` hands to Meca: the Mecables (fibers)
in registration order, their points, the interactions the emitters register
(sphere confinement from `Fiber::setInteractions`, crosslinker couples from
`Couple::setInteractions`), the simulation properties and, per step, the
N(0,1) Brownian draws.  Generation follows the reference's own placement rules:

* fiber length -> number of points by `FiberNaked::bestNbPoints`
  (fiber_naked.cc:22-30) and segment length fnCut = length/(points-1)
  (fiber_naked.cc:79-94);
* straight fibers centred on a random position inside the space with a random
  orientation (`setStraight(pos, dir, CENTER)`, fiber_naked.cc:57-76);
* variable lengths uniform in [min, max] (fiber_prop.cc:110-118 style);
* crosslinkers bound at uniformly random abscissas on pairs of fibers whose
  distance is within the binding range, attached to the nearest point of the
  partner fiber (couple_set.cc:185-235 attach semantics; an equilibrium
  binding state without running the Monte-Carlo stepping, which is out of
  scope);
* asters (Aster::build / buildSolid, aster.cc:121-300): a Solid made of a
  massive central bead of radius r and one drag-free point per fiber at
  distance r in a direction spread over the sphere, each fiber straight from
  the centre through its point (minus end at the centre), held by two links
  (Aster::setInteractions, aster.cc:53-89).

Everything here is host-side NumPy; nothing in this module touches the GPU.
"""
from dataclasses import dataclass, field

from typing import Optional

import numpy as np


def best_nb_points(ratio):
    """FiberNaked::bestNbPoints (fiber_naked.cc:22-30)."""
    n = int(ratio)
    if (2 * n + 1) * ratio > 2 * n * (n + 1):
        return n + 2
    return n + 1


@dataclass
class Scene:
    dim: int
    time_step: float
    kT: float
    viscosity: float
    hydro_radius0: float
    hydro_radius1: float
    nb_points: np.ndarray          # uint32 per fiber, registration order
    segmentation: np.ndarray       # fnCut per fiber
    rigidity: np.ndarray           # FiberProp::rigidity per fiber
    points: np.ndarray             # float64, DIM*sum(nb_points), matIndex order
    confine_radius: float = 0.0    # SpaceSphere radius (0: no confinement)
    confine_stiffness: float = 0.0
    # couples (Couple::setInteractions -> interLink(PI, PI)), call order
    link_fiber_a: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    link_abs_a: np.ndarray = field(default_factory=lambda: np.zeros(0))
    link_fiber_b: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    link_abs_b: np.ndarray = field(default_factory=lambda: np.zeros(0))
    link_stiffness: np.ndarray = field(default_factory=lambda: np.zeros(0))
    # per couple: 1 hand 1 attached, 2 hand 2 attached, 4 hand 1 is a motor, 8
    # hand 2 is a motor (None: attached passive crosslinkers); motor properties
    # (MotorProp max_speed, stall_force) of the motor hands
    link_state: Optional[np.ndarray] = None
    motor_max_speed: float = 0.0
    motor_stall_force: float = 0.0
    # Solids, registered after the fibers (simul_solve.cc:204-207): radius of
    # each point (0: no drag) and the initial reshape timer (solid.cc:28)
    sol_radius: list = field(default_factory=list)
    sol_timer: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    # aster links in call order, after the couples (simul_solve.cc:238-239):
    # kind 0 interLink(PE a1, PE b) (meca_inter.cc:196-230), kind 1
    # interLink(PI (a1, a2, ca), PE b) (meca_inter.cc:240-275)
    aster_kind: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    aster_a1: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    aster_a2: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    aster_ca: np.ndarray = field(default_factory=lambda: np.zeros(0))
    aster_b: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    aster_w: np.ndarray = field(default_factory=lambda: np.zeros(0))
    # the same links as (Solid point, fiber, abscissa) for the device emitter
    aster_s1: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    aster_s2: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    aster_fiber: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    aster_abs: np.ndarray = field(default_factory=lambda: np.zeros(0))
    seed: int = 0
    name: str = ""

    @property
    def nfib(self):
        return int(self.nb_points.size)

    @property
    def nsol(self):
        return len(self.sol_radius)

    @property
    def sol_nb_points(self):
        return np.array([r.size for r in self.sol_radius], dtype=np.uint32)

    @property
    def fiber_points_count(self):
        return int(self.nb_points.sum())

    @property
    def npts(self):
        return int(self.nb_points.sum()) + int(sum(r.size for r in self.sol_radius))

    @property
    def sol_index(self):
        """matIndex of each Solid (after every fiber point)."""
        n = self.sol_nb_points.astype(np.int64)
        return (self.fiber_points_count + np.concatenate([[0], np.cumsum(n)[:-1]])).astype(np.uint32)

    def solid_points(self, k):
        i = int(self.sol_index[k])
        return self.points[self.dim * i:self.dim * (i + self.sol_radius[k].size)]

    @property
    def naster_links(self):
        return int(self.aster_kind.size)

    @property
    def size(self):
        return self.dim * self.npts

    @property
    def mat_index(self):
        idx = np.zeros(self.nfib, dtype=np.uint32)
        np.cumsum(self.nb_points[:-1], out=idx[1:])
        return idx

    @property
    def nlinks(self):
        return int(self.link_fiber_a.size)

    def params(self):
        return dict(time_step=self.time_step, kT=self.kT, viscosity=self.viscosity,
                    hydro_radius0=self.hydro_radius0, hydro_radius1=self.hydro_radius1)

    def interpolate(self, fiber, abscissa):
        """FiberNaked::interpolate (fiber_naked.cc:1050-1065), vectorised;
        returns (mat1, mat2, coef)."""
        fiber = np.asarray(fiber, dtype=np.int64)
        ab = np.asarray(abscissa, dtype=np.float64)
        cut = self.segmentation[fiber]
        idx = self.mat_index.astype(np.int64)[fiber]
        last = self.nb_points.astype(np.int64)[fiber] - 1
        q = ab / cut
        n = np.floor(q)
        co = q - n
        rd = n.astype(np.int64)
        m1 = idx + rd
        coef = co.copy()
        beyond = rd >= last
        m1[beyond] = idx[beyond] + last[beyond] - 1
        coef[beyond] = 1.0
        neg = ab <= 0
        m1[neg] = idx[neg]
        coef[neg] = 0.0
        return m1.astype(np.uint32), (m1 + 1).astype(np.uint32), coef

    def gauss(self, step):
        """N(0,1) draws for one step (one per degree of freedom, object order)."""
        rng = np.random.default_rng([self.seed, 7919, step])
        return rng.standard_normal(self.size)


def _random_directions(rng, n, dim):
    if dim == 2:
        a = rng.uniform(0, 2 * np.pi, n)
        return np.stack([np.cos(a), np.sin(a)], axis=1)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _random_in_ball(rng, n, dim, radius):
    out = np.empty((0, dim))
    while out.shape[0] < n:
        c = rng.uniform(-radius, radius, (2 * n + 16, dim))
        c = c[(c * c).sum(axis=1) <= radius * radius]
        out = np.concatenate([out, c])
    return out[:n]


def _straight_fibers(rng, dim, lengths, seg, radius):
    nfib = lengths.size
    nbp = np.array([best_nb_points(L / seg) for L in lengths], dtype=np.uint32)
    cut = lengths / (nbp - 1)
    centers = _random_in_ball(rng, nfib, dim, radius)
    dirs = _random_directions(rng, nfib, dim)
    pts = []
    for f in range(nfib):
        start = centers[f] - 0.5 * dirs[f] * lengths[f]
        dp = dirs[f] * cut[f]
        p = np.arange(nbp[f])[:, None] * dp[None, :] + start[None, :]
        pts.append(p.ravel())
    return nbp, cut, np.concatenate(pts)


def _segment_segment(p0, d0, p1, d1):
    """Closest points between segments p0+s*d0 and p1+t*d1 (s,t in [0,1]),
    vectorised; returns s, t, distance."""
    r = p0 - p1
    a = (d0 * d0).sum(1)
    e = (d1 * d1).sum(1)
    b = (d0 * d1).sum(1)
    c = (d0 * r).sum(1)
    f = (d1 * r).sum(1)
    den = a * e - b * b
    s = np.where(den > 1e-14, np.clip((b * f - c * e) / np.where(den > 1e-14, den, 1), 0, 1), 0.0)
    t = (b * s + f) / e
    t0 = t < 0
    t1 = t > 1
    t = np.clip(t, 0, 1)
    s = np.where(t0, np.clip(-c / a, 0, 1), s)
    s = np.where(t1, np.clip((b - c) / a, 0, 1), s)
    q0 = p0 + s[:, None] * d0
    q1 = p1 + t[:, None] * d1
    return s, t, np.linalg.norm(q0 - q1, axis=1)


def _close_segment_pairs(dim, nbp, points, rng_range, return_params=False, group=None):
    """All pairs of segments of different fibers closer than rng_range (and
    of different groups when `group`, one label per fiber, is given)."""
    P = points.reshape(-1, dim)
    nbp = nbp.astype(np.int64)
    fib_of_pt = np.repeat(np.arange(nbp.size), nbp)
    starts = np.concatenate([[0], np.cumsum(nbp)[:-1]]).astype(np.int64)
    seg_pt = np.concatenate([s + np.arange(n - 1) for s, n in zip(starts, nbp)])
    seg_fib = fib_of_pt[seg_pt]
    A = P[seg_pt]
    B = P[seg_pt + 1]
    mid = 0.5 * (A + B)
    half = 0.5 * np.linalg.norm(B - A, axis=1).max()
    cell = 2 * half + rng_range
    key = np.floor(mid / cell).astype(np.int64)
    key -= key.min(axis=0)
    dims = key.max(axis=0) + 3
    def lin(k):
        out = np.zeros(k.shape[0], dtype=np.int64)
        for d in range(dim):
            out = out * dims[d] + (k[:, d] + 1)
        return out
    lk = lin(key)
    order = np.argsort(lk, kind="stable")
    slk = lk[order]
    pairs_i, pairs_j = [], []
    offs = np.array(np.meshgrid(*[[-1, 0, 1]] * dim, indexing="ij")).reshape(dim, -1).T
    for off in offs:
        nk = lin(key + off)
        lo = np.searchsorted(slk, nk, side="left")
        hi = np.searchsorted(slk, nk, side="right")
        cnt = hi - lo
        ii = np.repeat(np.arange(lk.size), cnt)
        within = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        jj = order[np.repeat(lo, cnt) + within]
        keep = (ii < jj) & (seg_fib[ii] != seg_fib[jj])
        if group is not None:
            keep &= group[seg_fib[ii]] != group[seg_fib[jj]]
        pairs_i.append(ii[keep])
        pairs_j.append(jj[keep])
    ii = np.concatenate(pairs_i)
    jj = np.concatenate(pairs_j)
    s, t, dist = _segment_segment(A[ii], B[ii] - A[ii], A[jj], B[jj] - A[jj])
    keep = dist <= rng_range
    if return_params:
        return seg_pt, seg_fib, ii[keep], jj[keep], s[keep], t[keep]
    return seg_pt, seg_fib, ii[keep], jj[keep]


def _bind_crosslinkers(rng, dim, nbp, cut, points, nlinks, binding_range):
    """Bind `nlinks` couples: hand 1 at a uniform abscissa on a fiber, hand 2 at
    the nearest point of another fiber within `binding_range`; attempts that
    find no partner are rejected (conditioned on segments with a partner)."""
    P = points.reshape(-1, dim)
    seg_pt, seg_fib, ci, cj = _close_segment_pairs(dim, nbp, points, binding_range)
    nbp = nbp.astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(nbp)[:-1]]).astype(np.int64)
    if ci.size == 0:
        raise ValueError("no fiber crossings within binding range")
    # each close pair seen from both sides
    a_seg = np.concatenate([ci, cj])
    b_seg = np.concatenate([cj, ci])
    cand = np.unique(a_seg)
    fa, aa, fb, ab = [], [], [], []
    while len(fa) < nlinks:
        m = 4 * (nlinks - len(fa)) + 64
        sa = cand[rng.integers(0, cand.size, m)]
        u = rng.random(m)
        pa = P[seg_pt[sa]] + u[:, None] * (P[seg_pt[sa] + 1] - P[seg_pt[sa]])
        for k in range(m):
            partners = b_seg[a_seg == sa[k]]
            q0 = P[seg_pt[partners]]
            d = P[seg_pt[partners] + 1] - q0
            t = np.clip(((pa[k] - q0) * d).sum(1) / (d * d).sum(1), 0, 1)
            dist = np.linalg.norm(q0 + t[:, None] * d - pa[k], axis=1)
            j = int(np.argmin(dist))
            if dist[j] <= binding_range:
                f1 = int(seg_fib[sa[k]])
                f2 = int(seg_fib[partners[j]])
                s1 = seg_pt[sa[k]] - starts[f1]
                s2 = seg_pt[partners[j]] - starts[f2]
                fa.append(f1)
                aa.append((s1 + u[k]) * cut[f1])
                fb.append(f2)
                ab.append((s2 + t[j]) * cut[f2])
                if len(fa) == nlinks:
                    break
    return (np.array(fa, np.uint32), np.array(aa), np.array(fb, np.uint32), np.array(ab))


def config1(seed=1):
    """2D, 20 fibers of length 11 (23 points), rigidity 20, inside a sphere of
    radius 5 (stiffness 200), viscosity 0.02, time_step 0.005, kT 0.0042."""
    rng = np.random.default_rng(seed)
    lengths = np.full(20, 11.0)
    nbp, cut, pts = _straight_fibers(rng, 2, lengths, 0.5, 5.0)
    return Scene(dim=2, time_step=0.005, kT=0.0042, viscosity=0.02, hydro_radius0=0.0125,
                 hydro_radius1=10.0, nb_points=nbp, segmentation=cut,
                 rigidity=np.full(20, 20.0), points=pts, confine_radius=5.0,
                 confine_stiffness=200.0, seed=seed, name="fibers-2d-20x23pt-sphere5")


def config2(seed=2, nfib=1000, nlinks=20000, radius=10.0, lmin=2.0, lmax=15.0):
    """3D, `nfib` fibers with lengths uniform in [lmin, lmax] (segmentation 0.5),
    rigidity 20, inside a sphere (stiffness 200), `nlinks` passive crosslinkers
    (stiffness 100, binding range 0.05), viscosity 0.02, time_step 0.005,
    kT 0.0042."""
    rng = np.random.default_rng(seed)
    lengths = rng.uniform(lmin, lmax, nfib)
    nbp, cut, pts = _straight_fibers(rng, 3, lengths, 0.5, radius)
    sc = Scene(dim=3, time_step=0.005, kT=0.0042, viscosity=0.02, hydro_radius0=0.0125,
               hydro_radius1=10.0, nb_points=nbp, segmentation=cut,
               rigidity=np.full(nfib, 20.0), points=pts, confine_radius=radius,
               confine_stiffness=200.0, seed=seed,
               name=f"fibers-3d-{nfib}-links-{nlinks}-sphere{radius:g}")
    if nlinks:
        fa, aa, fb, ab = _bind_crosslinkers(rng, 3, nbp, cut, pts, nlinks, 0.05)
        sc.link_fiber_a, sc.link_abs_a = fa, aa
        sc.link_fiber_b, sc.link_abs_b = fb, ab
        sc.link_stiffness = np.full(nlinks, 100.0)
    return sc


def mini(dim=3, seed=3, nfib=8, nlinks=20, radius=1.5, lmin=2.0, lmax=6.0, binding_range=0.5, lengths=None):
    """Small crosslinked scene for operator-level parity checks (fiber lengths
    uniform in [lmin, lmax], or the given lengths)."""
    rng = np.random.default_rng(seed)
    if lengths is None:
        lengths = rng.uniform(lmin, lmax, nfib)
    else:
        lengths = np.asarray(lengths, dtype=np.float64)
        nfib = lengths.size
    nbp, cut, pts = _straight_fibers(rng, dim, lengths, 0.5, radius)
    sc = Scene(dim=dim, time_step=0.005, kT=0.0042, viscosity=0.02, hydro_radius0=0.0125,
               hydro_radius1=10.0, nb_points=nbp, segmentation=cut,
               rigidity=np.full(nfib, 20.0), points=pts, confine_radius=radius,
               confine_stiffness=200.0, seed=seed, name=f"mini-{dim}d-{nfib}")
    if nlinks:
        fa, aa, fb, ab = _bind_crosslinkers(rng, dim, nbp, cut, pts, nlinks, binding_range)
        sc.link_fiber_a, sc.link_abs_a = fa, aa
        sc.link_fiber_b, sc.link_abs_b = fb, ab
        sc.link_stiffness = np.full(nlinks, 100.0)
    return sc


def _bind_crosslinkers_fast(rng, dim, nbp, cut, points, nlinks, binding_range, group=None):
    """Vectorised variant of _bind_crosslinkers for large scenes: hand 1 at a
    uniform abscissa within binding range of the closest approach of a random
    close segment pair, hand 2 at the nearest point of the partner segment;
    kept when the hands are within binding range."""
    P = points.reshape(-1, dim)
    seg_pt, seg_fib, ci, cj, cs, ct = _close_segment_pairs(dim, nbp, points, binding_range, True, group)
    if ci.size == 0:
        raise ValueError("no fiber crossings within binding range")
    nbp = nbp.astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(nbp)[:-1]]).astype(np.int64)
    out = [[], [], [], []]
    got = 0
    while got < nlinks:
        m = 3 * (nlinks - got) + 64
        k = rng.integers(0, ci.size, m)
        flip = rng.random(m) < 0.5
        sa = np.where(flip, cj[k], ci[k])
        sb = np.where(flip, ci[k], cj[k])
        s0 = np.where(flip, ct[k], cs[k])
        la = np.linalg.norm(P[seg_pt[sa] + 1] - P[seg_pt[sa]], axis=1)
        u = np.clip(s0 + rng.uniform(-1, 1, m) * binding_range / la, 0.0, 1.0)
        pa = P[seg_pt[sa]] + u[:, None] * (P[seg_pt[sa] + 1] - P[seg_pt[sa]])
        q0 = P[seg_pt[sb]]
        d = P[seg_pt[sb] + 1] - q0
        t = np.clip(((pa - q0) * d).sum(1) / (d * d).sum(1), 0.0, 1.0)
        dist = np.linalg.norm(q0 + t[:, None] * d - pa, axis=1)
        ok = np.nonzero(dist <= binding_range)[0][: nlinks - got]
        f1 = seg_fib[sa[ok]]
        f2 = seg_fib[sb[ok]]
        out[0].append(f1)
        out[1].append((seg_pt[sa[ok]] - starts[f1] + u[ok]) * cut[f1])
        out[2].append(f2)
        out[3].append((seg_pt[sb[ok]] - starts[f2] + t[ok]) * cut[f2])
        got += ok.size
    return (np.concatenate(out[0]).astype(np.uint32), np.concatenate(out[1]),
            np.concatenate(out[2]).astype(np.uint32), np.concatenate(out[3]))


def _sphere_directions(n, dim, rng=None):
    """n directions spread over the circle (2D, equal angles as aster.cc:273-282)
    or the sphere (3D, a Fibonacci lattice standing in for PointsOnSphere),
    randomly rotated when rng is given."""
    if dim == 2:
        a = 2 * np.pi * np.arange(n) / n + (rng.uniform(0, 2 * np.pi) if rng is not None else 0.0)
        return np.stack([np.cos(a), np.sin(a)], axis=1)
    i = np.arange(n)
    z = 1 - (2 * i + 1) / n
    r = np.sqrt(1 - z * z)
    phi = i * np.pi * (3 - np.sqrt(5))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    if rng is not None:
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        d = d @ q.T
    return d


def asters(dim=3, seed=3, nasters=10, fibers_per_aster=200, lmin=5.0, lmax=10.0, core=0.5, spread=4.0,
           radius=12.0, nmotors=50000, binding_range=0.05, stiffness=(1000.0, 500.0), motor_stiffness=100.0,
           min_separation=0.0, name=None):
    """Asters: each a Solid (central bead of radius `core` + one drag-free point
    per fiber at distance `core`) with `fibers_per_aster` fibers of lengths
    uniform in [lmin, lmax] (segmentation 0.5, rigidity 20) starting at the
    centre; aster centres uniform in a ball of radius `spread`; fibers confined
    in a sphere of radius `radius` (stiffness 200); `nmotors` couples
    (stiffness `motor_stiffness`) bound between fibers within `binding_range`
    (motor stepping frozen) between fibers of different asters (the fibers of
    one aster all meet at its centre, where every pair is within range);
    aster stiffness (clamp1 link, clamp2 link)."""
    rng = np.random.default_rng(seed)
    # aster centres at least `min_separation` apart (rejection sampling)
    centers = np.empty((0, dim))
    while centers.shape[0] < nasters:
        c = _random_in_ball(rng, 1, dim, spread)
        if centers.shape[0] == 0 or np.linalg.norm(centers - c, axis=1).min() >= min_separation:
            centers = np.concatenate([centers, c])
    nbp_l, cut_l, fpts, sol_r, sol_p, ends = [], [], [], [], [], []
    for a in range(nasters):
        dirs = _sphere_directions(fibers_per_aster, dim, rng)
        c = centers[a]
        sol_r.append(np.concatenate([[core], np.zeros(fibers_per_aster)]))
        sol_p.append(np.concatenate([c[None, :], c[None, :] + core * dirs]).ravel())
        lengths = rng.uniform(lmin, lmax, fibers_per_aster)
        for i in range(fibers_per_aster):
            n = best_nb_points(lengths[i] / 0.5)
            cut = lengths[i] / (n - 1)
            nbp_l.append(n)
            cut_l.append(cut)
            fpts.append((c[None, :] + np.arange(n)[:, None] * (cut * dirs[i])[None, :]).ravel())
    nbp = np.array(nbp_l, dtype=np.uint32)
    cut = np.array(cut_l)
    fiber_pts = np.concatenate(fpts)
    nfib = nbp.size
    sc = Scene(dim=dim, time_step=0.005, kT=0.0042, viscosity=0.02, hydro_radius0=0.0125,
               hydro_radius1=10.0, nb_points=nbp, segmentation=cut, rigidity=np.full(nfib, 20.0),
               points=np.concatenate([fiber_pts] + sol_p), confine_radius=radius, confine_stiffness=200.0,
               seed=seed, name=name or f"asters-{dim}d-{nasters}x{fibers_per_aster}-motors-{nmotors}")
    sc.sol_radius = sol_r
    sc.sol_timer = rng.integers(0, 7, nasters)
    if nmotors:
        group = np.repeat(np.arange(nasters), fibers_per_aster)
        fa, aa, fb, ab = _bind_crosslinkers_fast(rng, dim, nbp, cut, fiber_pts, nmotors, binding_range, group)
        sc.link_fiber_a, sc.link_abs_a, sc.link_fiber_b, sc.link_abs_b = fa, aa, fb, ab
        sc.link_stiffness = np.full(fa.size, motor_stiffness)
        # both hands of every couple are motors (speed 0.5 um/s, stall force 5 pN)
        sc.link_state = np.full(fa.size, 15, np.uint8)
        sc.motor_max_speed = 0.5
        sc.motor_stall_force = 5.0
    # Aster::setInteractions (aster.cc:53-89): per fiber, centre -- minus end,
    # then, for a fiber longer than the clamp distance `core`, the fiber at
    # abscissa `core` -- its Solid point (kind 1), otherwise the plus end -- the
    # Solid's centre/point interpolated at length/core (kind 2)
    fidx = sc.mat_index
    sidx = sc.sol_index
    kind, a1, a2, ca, b, w, s1, s2, fl, ab = [], [], [], [], [], [], [], [], [], []
    fib = np.arange(nfib)
    m1, m2, co = sc.interpolate(fib, np.full(nfib, core))
    for a in range(nasters):
        for i in range(fibers_per_aster):
            f = a * fibers_per_aster + i
            length = cut[f] * (int(nbp[f]) - 1)      # FiberNaked::length()
            clamp2 = sidx[a] + 1 + i
            if length > core:
                kind += [0, 1]
                a1 += [sidx[a], m1[f]]
                a2 += [sidx[a], m2[f]]
                ca += [0.0, co[f]]
                b += [fidx[f], clamp2]
                s2 += [sidx[a], clamp2]
                ab += [0.0, core]
            else:
                c = length / core
                kind += [0, 2]
                a1 += [sidx[a], sidx[a]]
                a2 += [sidx[a], clamp2]
                ca += [0.0, c]
                b += [fidx[f], fidx[f] + int(nbp[f]) - 1]
                s2 += [sidx[a], clamp2]
                ab += [0.0, c]
            w += [stiffness[0], stiffness[1]]
            s1 += [sidx[a], sidx[a] if length <= core else clamp2]
            fl += [f, f]
    sc.aster_s1 = np.array(s1, np.uint32)
    sc.aster_s2 = np.array(s2, np.uint32)
    sc.aster_fiber = np.array(fl, np.uint32)
    sc.aster_abs = np.array(ab)
    sc.aster_kind = np.array(kind, np.int64)
    sc.aster_a1 = np.array(a1, np.uint32)
    sc.aster_a2 = np.array(a2, np.uint32)
    sc.aster_ca = np.array(ca)
    sc.aster_b = np.array(b, np.uint32)
    sc.aster_w = np.array(w)
    return sc


def config3(seed=3):
    """10 asters x 200 fibers (lengths 5-10 um), 50000 motors (stiffness 100),
    aster stiffness (1000, 500), in a sphere of radius 12, aster centres at
    least 3 um apart within 5 um of the centre."""
    return asters(dim=3, seed=seed, spread=5.0, min_separation=3.0)


def mini_asters(dim=3, seed=11, nasters=2, fibers_per_aster=6, nmotors=12, lmin=1.5, lmax=3.0, core=0.5):
    """Small aster scene for operator-level parity checks; with lmin below
    `core` some fibers are shorter than the clamp distance (aster.cc:77-84)."""
    return asters(dim=dim, seed=seed, nasters=nasters, fibers_per_aster=fibers_per_aster, lmin=lmin, lmax=lmax,
                  core=core, spread=1.0, radius=2.5, nmotors=nmotors, binding_range=0.3,
                  name=f"mini-asters-{dim}d-{nasters}x{fibers_per_aster}")


@dataclass
class Batch:
    """Independent systems solved together: one Scene holding every fiber, and
    the first fiber of each system."""
    scene: Scene
    first_fiber: np.ndarray
    systems: list

    @property
    def nsys(self):
        return int(self.first_fiber.size)


def concatenate(systems, name="batch"):
    """Stack independent scenes (same dim and properties) into one Batch."""
    s0 = systems[0]
    nfib = np.array([sc.nfib for sc in systems])
    first = np.concatenate([[0], np.cumsum(nfib)[:-1]]).astype(np.uint32)
    cat = lambda attr: np.concatenate([getattr(sc, attr) for sc in systems])
    big = Scene(dim=s0.dim, time_step=s0.time_step, kT=s0.kT, viscosity=s0.viscosity,
                hydro_radius0=s0.hydro_radius0, hydro_radius1=s0.hydro_radius1,
                nb_points=cat("nb_points").astype(np.uint32), segmentation=cat("segmentation"),
                rigidity=cat("rigidity"), points=cat("points"),
                confine_radius=s0.confine_radius, confine_stiffness=s0.confine_stiffness,
                link_fiber_a=np.concatenate([sc.link_fiber_a + f for sc, f in zip(systems, first)]).astype(np.uint32),
                link_abs_a=cat("link_abs_a"),
                link_fiber_b=np.concatenate([sc.link_fiber_b + f for sc, f in zip(systems, first)]).astype(np.uint32),
                link_abs_b=cat("link_abs_b"), link_stiffness=cat("link_stiffness"),
                seed=s0.seed, name=name)
    big.gauss = lambda step: np.concatenate([sc.gauss(step) for sc in systems])
    return Batch(scene=big, first_fiber=first, systems=systems)


def config5(seed=5, nsys=64, fmin=50, fmax=2000, links_min=5, links_max=30, radius=10.0):
    """64 independent 3D systems: fiber counts uniform in [fmin, fmax], lengths
    uniform in [2, 15] um, 5-30 crosslinkers per fiber, other properties as the
    1000-fiber crosslinked scene."""
    rng = np.random.default_rng(seed)
    systems = []
    for k in range(nsys):
        nf = int(rng.integers(fmin, fmax + 1))
        srng = np.random.default_rng([seed, k])
        lengths = srng.uniform(2.0, 15.0, nf)
        nbp, cut, pts = _straight_fibers(srng, 3, lengths, 0.5, radius)
        nl = int(srng.integers(links_min, links_max + 1)) * nf
        sc = Scene(dim=3, time_step=0.005, kT=0.0042, viscosity=0.02, hydro_radius0=0.0125,
                   hydro_radius1=10.0, nb_points=nbp, segmentation=cut, rigidity=np.full(nf, 20.0),
                   points=pts, confine_radius=radius, confine_stiffness=200.0, seed=seed * 1000 + k,
                   name=f"system-{k}")
        try:
            fa, aa, fb, ab = _bind_crosslinkers_fast(srng, 3, nbp, cut, pts, nl, 0.05)
        except ValueError:
            # a sparse system may have no fiber crossing within binding range
            fa, aa, fb, ab = (np.zeros(0, np.uint32), np.zeros(0), np.zeros(0, np.uint32), np.zeros(0))
        sc.link_fiber_a, sc.link_abs_a, sc.link_fiber_b, sc.link_abs_b = fa, aa, fb, ab
        sc.link_stiffness = np.full(fa.size, 100.0)
        systems.append(sc)
    return concatenate(systems, name=f"batch-{nsys}-systems")


def mini_batch(seed=7, nsys=4, dim=3):
    """Small batch of independent crosslinked scenes for parity checks."""
    systems = [mini(dim=dim, seed=seed + 13 * k, nfib=5 + 3 * k, nlinks=10 + 5 * k) for k in range(nsys)]
    return concatenate(systems, name=f"mini-batch-{nsys}")
