"""The interactions that read the positions (meca_inter.cc), evaluated on the
device from the prepared points, against the oracle: indices and values of
mB, mC and vBAS bit-exact, calls of every kind interleaved with host-recorded
ones so that many touch the same entries and the call order decides the sums.
The torque's angle comes from the device atan2, within an ulp or two of the
host's: its entries are compared to 1e-13 relative."""
import numpy as np
import pytest

from conftest import oracle_from_scene
from test_gpu_parity import gpu_from_scene, nr_entries

pytestmark = pytest.mark.gpu


def _segments(sc):
    idx = np.zeros(sc.nfib, np.int64)
    np.cumsum(sc.nb_points[:-1], out=idx[1:])
    return np.array([int(i) + k for i, n in zip(idx, sc.nb_points) for k in range(int(n) - 1)])


def _entries(gpu, orc, which):
    r, c, v = gpu.matrix_entries(which)
    rr, cc, vv = nr_entries(orc, which)
    keep = np.ones(rr.size, bool)
    touched = set(c.tolist())
    for i in range(rr.size):
        if rr[i] == cc[i] and cc[i] not in touched and vv[i] == 0.0:
            keep[i] = False
    rr, cc, vv = rr[keep], cc[keep], vv[keep]
    o1, o2 = np.lexsort((r, c)), np.lexsort((rr, cc))
    return (r[o1], c[o1], v[o1]), (rr[o2], cc[o2], vv[o2])


def _assert_same(pkg, oracle, gpu, orc, rtol=0.0):
    for which in (0, 1):
        (r, c, v), (rr, cc, vv) = _entries(gpu, orc, which)
        assert np.array_equal(r, rr) and np.array_equal(c, cc), f"matrix {which}: index sets differ"
        if rtol == 0.0:
            assert np.array_equal(v, vv), f"matrix {which}: values must be bit-exact"
        else:
            assert np.abs(v - vv).max() <= rtol * max(np.abs(vv).max(), 1e-300)
    b, bb = gpu.vector(pkg.VEC_BAS), orc.vector(oracle.BAS)
    if rtol == 0.0:
        assert np.array_equal(b, bb)
    else:
        assert np.abs(b - bb).max() <= rtol * max(np.abs(bb).max(), 1e-300)


def _emit(m, sc, rng, dim, ncalls):
    """the same sequence of calls on either side (the rng is seeded alike)"""
    segs = _segments(sc)
    npts = sc.npts
    pts = np.asarray(sc.points).reshape(-1, dim)
    for k in range(ncalls):
        s1, s2 = (int(v) for v in rng.choice(segs, 2, replace=False))
        b = int(rng.integers(0, npts))
        ca, cb = float(rng.uniform(0.05, 0.95)), float(rng.uniform(0.05, 0.95))
        w = float(rng.uniform(1.0, 100.0))
        length = float(rng.uniform(0.05, 1.5))
        g = rng.standard_normal(dim)
        nv = rng.standard_normal(dim)
        nv /= np.linalg.norm(nv)
        kind = k % 17
        if kind == 0:
            m.inter_long_link_ie(s1, s1 + 1, ca, b, length, w)
        elif kind == 1:
            m.inter_long_link_ii(s1, s1 + 1, ca, s2, s2 + 1, cb, length, w)
        elif kind == 2:
            m.inter_side_link_ie(s1, s1 + 1, ca, b, length, w)
        elif kind == 3:
            m.inter_side_link_ii(s1, s1 + 1, ca, s2, s2 + 1, cb, length, w)
        elif kind == 4:
            if dim == 2:
                m.inter_side_side_link(s1, s1 + 1, ca, s2, s2 + 1, cb, length, w)
        elif kind == 5:
            m.inter_sliding_link_ie(s1, s1 + 1, ca, b, w)
        elif kind == 6:
            m.inter_sliding_link_ii(s1, s1 + 1, ca, s2, s2 + 1, cb, w)
        elif kind == 7:
            m.inter_side_sliding_link_ie(s1, s1 + 1, ca, b, length, w)
        elif kind == 8:
            m.inter_side_sliding_link_ii(s1, s1 + 1, ca, s2, s2 + 1, cb, length, w)
        elif kind == 9:
            m.inter_side_clamp(s1, s1 + 1, ca, g, length, w)
        elif kind == 10:
            m.inter_sliding_clamp(s1, s1 + 1, ca, g, w)
        elif kind == 11:
            m.inter_plane_i(s1, s1 + 1, ca, nv, g, w)
        elif kind == 12:
            m.add_pure_force_i(s1, s1 + 1, ca, g)
        elif kind == 13:
            m.inter_long_link_ee(b, int(rng.integers(0, npts)), length, w)
        elif kind == 14:
            # at its own center the long clamp records nothing
            center = pts[b] if rng.uniform() < 0.3 else g
            m.inter_long_clamp(b, center, length, w)
        elif kind == 15:
            m.inter_link_ie(s1, s1 + 1, ca, b, w)
        else:
            # overlapping points: nothing recorded
            m.inter_side_link_ie(s1, s1 + 1, ca, s1 + 1, length, w)
            m.inter_sliding_link_ie(s1, s1 + 1, ca, s1, w)


@pytest.mark.parametrize("dim", [2, 3])
def test_position_emitters_bit_exact(pkg, scenes, oracle, dim):
    sc = scenes.mini(dim=dim, seed=23, nfib=6, nlinks=10)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    for m in (gpu, orc):
        _emit(m, sc, np.random.default_rng(300 + dim), dim, 170)
    g0 = sc.gauss(0)
    gpu.setup(g0)
    orc.setup(g0)
    _assert_same(pkg, oracle, gpu, orc)
    # a second step from moved points: the calls read the new positions
    p1 = np.asarray(sc.points) + 0.05 * np.random.default_rng(9).standard_normal(sc.size)
    sc.points = p1
    gpu.prepare(p1)
    orc.prepare(p1)
    for m in (gpu, orc):
        _emit(m, sc, np.random.default_rng(400 + dim), dim, 120)
    gpu.setup(g0)
    orc.setup(g0)
    _assert_same(pkg, oracle, gpu, orc)


def test_torque(pkg, scenes, oracle):
    sc = scenes.mini(dim=2, seed=29, nfib=6, nlinks=6)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    segs = _segments(sc)
    for m in (gpu, orc):
        rng = np.random.default_rng(77)
        for k in range(60):
            s1, s2 = (int(v) for v in rng.choice(segs, 2, replace=False))
            ang = float(rng.uniform(-3.0, 3.0))
            m.inter_torque(s1, s1 + 1, 0.5, s2, s2 + 1, 0.5, np.cos(ang), np.sin(ang),
                           float(rng.uniform(1.0, 50.0)))
            if k % 3 == 0:
                m.inter_link_ee(s1, s2, 5.0)
    g0 = sc.gauss(0)
    gpu.setup(g0)
    orc.setup(g0)
    _assert_same(pkg, oracle, gpu, orc, rtol=1e-13)


def test_plane_only_emitters_refused_in_3d(pkg, scenes):
    sc = scenes.mini(dim=3, seed=23, nfib=4, nlinks=4)
    gpu = gpu_from_scene(pkg, sc)
    segs = _segments(sc)
    s1, s2 = int(segs[0]), int(segs[-1])
    with pytest.raises(pkg.CymgError):
        gpu.inter_torque(s1, s1 + 1, 0.5, s2, s2 + 1, 0.5, 1.0, 0.0, 1.0)
    with pytest.raises(pkg.CymgError):
        gpu.inter_side_side_link(s1, s1 + 1, 0.5, s2, s2 + 1, 0.5, 0.2, 1.0)
    # an interpolated point across two fibers is refused
    idx = int(sc.nb_points[0])
    with pytest.raises(pkg.CymgError):
        gpu.inter_sliding_clamp(idx - 1, idx, 0.5, np.zeros(3), 1.0)
