"""Segments near places on the device (cymg_nearby_segments: FiberGrid::
nearbySegments / tryToAttach's range test, fiber_grid.cc:248-337) against the
oracle's brute force on the same points: identical counts, fibers, points,
abscissas and squared distances (bit for bit: the same arithmetic, no FMA
contraction), in 2D and 3D, strict and inclusive, with an excluded fiber per
place, at several ranges, for places near fibers and anywhere in the space,
and at the 1000-fiber size."""
import numpy as np
import pytest

from conftest import oracle_from_scene
from test_gpu_parity import gpu_from_scene

pytestmark = pytest.mark.gpu


def _places(sc, rng, n):
    """Half near fiber points (within 0.3), half uniform in the bounding box."""
    P = sc.points.reshape(-1, sc.dim)
    near = P[rng.integers(0, P.shape[0], n // 2)] + rng.uniform(-0.3, 0.3, (n // 2, sc.dim))
    lo, hi = P.min(0) - 0.5, P.max(0) + 0.5
    anywhere = rng.uniform(lo, hi, (n - n // 2, sc.dim))
    return np.concatenate([near, anywhere])


def _compare(gpu, orc, W, R, inclusive, exclude=None):
    counts, (f, p, a, d) = gpu.nearby_segments(W, R, inclusive=inclusive, exclude=exclude)
    o = 0
    for i in range(W.shape[0]):
        ex = -1 if exclude is None else int(exclude[i])
        of, op, oa, od = orc.nearby_segments(W[i], R, inclusive=inclusive, exclude=ex)
        n = int(counts[i])
        assert n == of.size, i
        assert np.array_equal(f[o:o + n], of) and np.array_equal(p[o:o + n], op), i
        assert np.array_equal(a[o:o + n], oa) and np.array_equal(d[o:o + n], od), i
        o += n
    assert o == f.size
    return f.size


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("inclusive", [False, True])
def test_nearby_segments_match_oracle(pkg, scenes, oracle, dim, inclusive):
    print(f"grid: dim={dim} inclusive={inclusive}", flush=True)
    sc = scenes.mini(dim=dim, seed=41 + dim, nfib=20, nlinks=30, radius=2.5)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    rng = np.random.default_rng(dim)
    W = _places(sc, rng, 400)
    kept = 0
    for R in (0.05, 0.2, 0.6):
        kept += _compare(gpu, orc, W, R, inclusive)
    ex = rng.integers(0, sc.nfib, W.shape[0]).astype(np.uint32)
    kept += _compare(gpu, orc, W, 0.2, inclusive, exclude=ex)
    assert kept > 100
    gpu.close()


def test_nearby_segments_full_size(pkg, scenes, oracle):
    sc = scenes.config2(seed=2)
    print(f"grid: full size {sc.name}", flush=True)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    W = _places(sc, np.random.default_rng(7), 2000)
    assert _compare(gpu, orc, W, 0.05, True) > 10
    assert _compare(gpu, orc, W, 0.25, False) > 200
    gpu.close()


def test_nearby_segments_with_solids(pkg, scenes, oracle):
    """Solids (asters) have no segments: only the fibers are listed."""
    print("grid: asters", flush=True)
    sc = scenes.mini_asters(dim=3, seed=12)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    W = _places(sc, np.random.default_rng(5), 300)
    assert _compare(gpu, orc, W, 0.3, False) > 20
    gpu.close()


def test_nearby_segments_validation(pkg, scenes):
    sc = scenes.mini(dim=3, seed=3, nfib=4, nlinks=5)
    gpu = gpu_from_scene(pkg, sc)
    with pytest.raises(pkg.CymgError):
        gpu.nearby_segments(np.zeros((1, 3)), -1.0)
    counts, out = gpu.nearby_segments(np.zeros((0, 3)), 0.1)
    assert counts.size == 0 and out[0].size == 0
    gpu.close()


def test_more_than_64_segments_near_a_place_is_refused(pkg, scenes, oracle):
    """A range that brings more than 64 segments near one place: the call
    fails with CYMG_ERR_UNSUPPORTED before writing any output."""
    sc = scenes.config2(seed=47, nfib=80, nlinks=50, radius=2.5)
    gpu = gpu_from_scene(pkg, sc)
    orc = oracle_from_scene(oracle, sc)
    W = np.zeros((2, 3))
    assert orc.nearby_segments(W[0], 2.0)[0].size > 64
    import ctypes as C
    counts = np.full(2, 7, np.uint32)
    total = C.c_uint32(12345)
    print("grid: overflow refused", flush=True)
    rc = gpu.L.cymg_nearby_segments(gpu.h, 2, W.ctypes.data_as(C.c_void_p), None, 2.0, 0,
                                    counts.ctypes.data_as(C.c_void_p), 0, None, None, None, None, C.byref(total))
    assert rc == pkg.ERR_UNSUPPORTED
    assert np.all(counts == 7) and total.value == 12345
    gpu.close()
