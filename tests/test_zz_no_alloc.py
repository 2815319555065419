"""The no-allocation FiberGrid query (cymg_nearby_segments_noalloc) through
the C ABI against the allocating one (cymg_nearby_segments) on the same
points: identical counts, fibers, points, abscissas and squared distances, on
two calls in a row, the second with fewer places (the shared array keeps the
first call's contents, which must not leak into the second), then a size
beyond the array refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_parity import gpu_from_scene

pytestmark = pytest.mark.gpu


def _query(gpu, fn, W, R, inclusive, exclude, cap):
    nq = W.shape[0]
    W = np.ascontiguousarray(W, dtype=np.float64)
    ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.uint32)
    counts = np.zeros(nq, np.uint32)
    f, p = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
    a, d = np.zeros(cap), np.zeros(cap)
    total = C.c_uint32(0)
    vp = lambda x: None if x is None else x.ctypes.data_as(C.c_void_p)
    rc = fn(gpu.h, nq, vp(W), vp(ex), float(R), int(inclusive), vp(counts), cap, vp(f), vp(p), vp(a), vp(d),
            C.byref(total))
    return rc, counts, total.value, (f[:total.value], p[:total.value], a[:total.value], d[:total.value])


def test_noalloc_query_matches_the_allocating_one(pkg, scenes):
    sc = scenes.config2(seed=3, nfib=300, nlinks=1000, radius=6.0)
    gpu = gpu_from_scene(pkg, sc)
    rng = np.random.default_rng(9)
    P = sc.points.reshape(-1, 3)
    big = P[rng.integers(0, P.shape[0], 20000)] + rng.uniform(-0.2, 0.2, (20000, 3))
    small = P[rng.integers(0, P.shape[0], 300)] + rng.uniform(-0.2, 0.2, (300, 3))
    ex_small = rng.integers(0, sc.nfib, 300).astype(np.uint32)
    for W, R, inclusive, ex in ((big, 0.12, False, None), (small, 0.08, True, ex_small)):
        print(f"no-alloc query: places={W.shape[0]} range={R} inclusive={inclusive}", flush=True)
        cap = 40 * W.shape[0]
        r1, c1, t1, o1 = _query(gpu, gpu.L.cymg_nearby_segments, W, R, inclusive, ex, cap)
        r2, c2, t2, o2 = _query(gpu, gpu.L.cymg_nearby_segments_noalloc, W, R, inclusive, ex, cap)
        assert r1 == 0 and r2 == 0, gpu.L.cymg_last_error(gpu.h)
        assert t1 == t2 > 0 and t1 <= cap
        assert np.array_equal(c1, c2)
        for x, y in zip(o1, o2):
            assert np.array_equal(x, y)
    # beyond the array: refused with a clear error, before any GPU work
    W = np.zeros((65537, 3))
    r, *_ = _query(gpu, gpu.L.cymg_nearby_segments_noalloc, W, 0.1, False, None, 16)
    assert r == pkg.ERR_UNSUPPORTED
    assert b"CYMG_NOALLOC_MAX_PLACES" in gpu.L.cymg_last_error(gpu.h)
    gpu.close()
