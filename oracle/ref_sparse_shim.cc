// C-ABI driver of the reference's own MatrixSparseSymmetric1
// (src/math/matsparsesym1.{h,cc}), compiled from the reference sources where
// they lie by `make ref` in this folder, into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/test_ref_sparse.py to check the
// oracle's restatement of the mB/mC storage (element creation order, the
// Numerical-Recipes layout, vecMulAdd / vecMulAddIso2D / vecMulAddIso3D and
// the block extraction used by Meca::getBlock) against the reference's code.
// Only the class's public interface is used.
#include "matsparsesym1.h"

extern "C" {

void* rsp_create(unsigned n)
{
    MatrixSparseSymmetric1* m = new MatrixSparseSymmetric1();
    m->allocate(n);
    m->makeZero();
    return m;
}

void rsp_destroy(void* h) { delete static_cast<MatrixSparseSymmetric1*>(h); }

void rsp_zero(void* h) { static_cast<MatrixSparseSymmetric1*>(h)->makeZero(); }

// M(i, j) += v, through the element reference the emitters write into
void rsp_add(void* h, unsigned i, unsigned j, double v) { (*static_cast<MatrixSparseSymmetric1*>(h))(i, j) += v; }

unsigned rsp_nnz(const void* h) { return static_cast<const MatrixSparseSymmetric1*>(h)->nbNonZeroElements(); }

void rsp_prepare(void* h) { static_cast<MatrixSparseSymmetric1*>(h)->prepareForMultiply(); }

int rsp_nonzero(const void* h) { return static_cast<const MatrixSparseSymmetric1*>(h)->nonZero() ? 1 : 0; }

// Y += M X: dim 1 = vecMulAdd, 2 = vecMulAddIso2D, 3 = vecMulAddIso3D
void rsp_mul_add(const void* h, int dim, const double* X, double* Y)
{
    const MatrixSparseSymmetric1* m = static_cast<const MatrixSparseSymmetric1*>(h);
    if (dim == 3)
        m->vecMulAddIso3D(X, Y);
    else if (dim == 2)
        m->vecMulAddIso2D(X, Y);
    else
        m->vecMulAdd(X, Y);
}

// M += the diagonal block [x, x+sx) (full = addDiagonalBlock, else addTriangularBlock)
void rsp_block(const void* h, int full, unsigned x, unsigned sx, double* M)
{
    const MatrixSparseSymmetric1* m = static_cast<const MatrixSparseSymmetric1*>(h);
    if (full)
        m->addDiagonalBlock(M, x, sx);
    else
        m->addTriangularBlock(M, x, sx);
}

}  // extern "C"
