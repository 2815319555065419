"""The oracle's mB/mC storage against the reference's own MatrixSparseSymmetric1
(src/math/matsparsesym1.cc, built from its sources by `make -C oracle ref`
into oracle/_ref/): the same sequences of `M(i,j) += v` give the same number
of stored elements and bit-identical products (vecMulAdd, vecMulAddIso2D,
vecMulAddIso3D: the element order inside each column fixes the rounding) and
diagonal blocks (addDiagonalBlock, addTriangularBlock, as Meca::getBlock
reads them). Skipped where the reference library has not been built."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po

REF_LIB = os.path.join(os.path.dirname(po.__file__), "_ref", "libref_sparse.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (needs the reference sources)")

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _bind():
    R = C.CDLL(REF_LIB)
    R.rsp_create.restype = C.c_void_p
    R.rsp_create.argtypes = [C.c_uint]
    for f in ("rsp_destroy", "rsp_zero", "rsp_prepare"):
        getattr(R, f).argtypes = [C.c_void_p]
    R.rsp_add.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_double]
    R.rsp_nnz.restype = C.c_uint
    R.rsp_nnz.argtypes = [C.c_void_p]
    R.rsp_nonzero.argtypes = [C.c_void_p]
    R.rsp_mul_add.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p]
    R.rsp_block.argtypes = [C.c_void_p, C.c_int, C.c_uint, C.c_uint, _f64p]
    O = po.lib()
    O.orc_sparse_create.restype = C.c_void_p
    O.orc_sparse_create.argtypes = [C.c_uint32]
    for f in ("orc_sparse_destroy", "orc_sparse_zero", "orc_sparse_prepare"):
        getattr(O, f).argtypes = [C.c_void_p]
    O.orc_sparse_add.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double]
    O.orc_sparse_nnz.restype = C.c_uint32
    O.orc_sparse_nnz.argtypes = [C.c_void_p]
    O.orc_sparse_nonzero.argtypes = [C.c_void_p]
    O.orc_sparse_mul_add.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p]
    O.orc_sparse_block.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, _f64p]
    return R, O


class Pair:
    """The reference matrix and the oracle's, driven by the same calls."""

    def __init__(self, n):
        self.R, self.O = _bind()
        self.n = n
        self.r = self.R.rsp_create(n)
        self.o = self.O.orc_sparse_create(n)

    def close(self):
        self.R.rsp_destroy(self.r)
        self.O.orc_sparse_destroy(self.o)

    def zero(self):
        self.R.rsp_zero(self.r)
        self.O.orc_sparse_zero(self.o)

    def add(self, i, j, v):
        self.R.rsp_add(self.r, int(i), int(j), float(v))
        self.O.orc_sparse_add(self.o, int(i), int(j), float(v))

    def check(self, rng, blocks=8):
        R, O, n = self.R, self.O, self.n
        assert R.rsp_nnz(self.r) == O.orc_sparse_nnz(self.o)
        assert R.rsp_nonzero(self.r) == O.orc_sparse_nonzero(self.o)
        R.rsp_prepare(self.r)
        O.orc_sparse_prepare(self.o)
        for dim in (1, 2, 3):
            X = rng.standard_normal(dim * n)
            Y0 = rng.standard_normal(dim * n)
            yr, yo = Y0.copy(), Y0.copy()
            R.rsp_mul_add(self.r, dim, X, yr)
            O.orc_sparse_mul_add(self.o, dim, X, yo)
            assert np.array_equal(yr, yo), dim
        for _ in range(blocks):
            x = int(rng.integers(0, n))
            sx = int(rng.integers(1, n - x + 1))
            for full in (0, 1):
                M0 = rng.standard_normal(sx * sx)
                mr, mo = M0.copy(), M0.copy()
                R.rsp_block(self.r, full, x, sx, mr)
                O.orc_sparse_block(self.o, full, x, sx, mo)
                assert np.array_equal(mr, mo), (x, sx, full)


@pytest.mark.parametrize("n,writes", [(1, 5), (7, 40), (60, 900), (500, 6000)])
def test_random_writes_match_reference(n, writes):
    """Random (i, j) in any order (both triangles, the diagonal, repeats):
    exercises the diagonal-first columns, insertion in row order and the
    capacity growth in chunks of 4."""
    rng = np.random.default_rng(n)
    p = Pair(n)
    for _ in range(writes):
        p.add(rng.integers(0, n), rng.integers(0, n), rng.standard_normal())
    p.check(rng)
    p.close()


def test_link_shaped_writes_and_reuse_after_zero():
    """Writes shaped like Meca::interLink (the upper triangle of -w c c' over
    the two or four points of a link, meca_inter.cc:285-321), then makeZero
    and a different pattern in the same allocation, as from one step to the
    next."""
    rng = np.random.default_rng(7)
    n = 300
    p = Pair(n)
    for step in range(3):
        p.zero()
        for _ in range(400):
            a = int(rng.integers(0, n - 1))
            b = int(rng.integers(0, n - 1))
            ca, cb, w = rng.random(), rng.random(), 50 + 100 * rng.random()
            idx = [a, a + 1, b, b + 1]
            c = np.array([1 - ca, ca, -(1 - cb), -cb])
            for x in range(4):
                for y in range(x, 4):
                    p.add(idx[x], idx[y], -w * c[x] * c[y])
        p.check(rng)
    p.close()


def test_empty_and_explicit_zeros():
    rng = np.random.default_rng(3)
    p = Pair(20)
    p.check(rng)                       # nothing stored: products leave Y as it is
    for k in range(10):
        p.add(k, 19 - k, 0.0)          # stored elements that hold zeros
    p.check(rng)
    p.close()
