// Device-side application of one preconditioner block: y = Binv x
// (blas_xgemv('N') on a column-major block, meca.cc:458), or a copy when the
// block could not be factorised (meca.cc:463).  One wavefront per fiber; lane
// i owns rows i, i+64, ...; consecutive lanes read consecutive addresses of a
// column.  T is the storage type of the inverse: double (FP64 mode) or float
// (FP32 mode: factorised in FP32, stored in FP32, half the HBM bytes).  The
// operand and the accumulation stay FP64 in both modes, so the FP32 mode
// applies a fixed linear operator: rounding the operand to FP32 as well makes
// the preconditioner nonlinear and multiplied BiCGStab iterations by ~9 on the
// crosslinked 300-fiber scene (1225 vs 137 iterations at 1e-6), while FP64
// accumulation over FP32 entries converges like the FP64 blocks.
#pragma once

#include "common.h"

namespace cymg {

// xs: the operand, already converted to T, in LDS; xd: the same operand in
// FP64 (for the copy case); y: output (HBM or LDS).
template <typename T>
__device__ __forceinline__ void block_apply(const T* __restrict__ B, int bs, bool use, const double* xs,
                                            const double* xd, double* y)
{
    for (int i = threadIdx.x; i < bs; i += kWave) {
        if (!use) {
            y[i] = xd[i];
            continue;
        }
        double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        const T* col = B + i;
        int j = 0;
        for (; j + 3 < bs; j += 4) {
            a0 += (double)col[(size_t)bs * j] * xs[j];
            a1 += (double)col[(size_t)bs * (j + 1)] * xs[j + 1];
            a2 += (double)col[(size_t)bs * (j + 2)] * xs[j + 2];
            a3 += (double)col[(size_t)bs * (j + 3)] * xs[j + 3];
        }
        for (; j < bs; ++j) a0 += (double)col[(size_t)bs * j] * xs[j];
        y[i] = (a0 + a1) + (a2 + a3);
    }
}

}  // namespace cymg
