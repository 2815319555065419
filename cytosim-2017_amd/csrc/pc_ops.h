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

// Same product with NW wavefronts per fiber: wave w takes the columns
// j = w, w+NW, ..., lane i the rows i, i+64, ...; the NW partial sums are
// added in wave order through LDS (red: NW*bs doubles), so the result does not
// depend on scheduling.  More wavefronts per block keep more loads in flight.
template <typename T, int NW>
__device__ __forceinline__ void block_apply_mw(const T* __restrict__ B, int bs, bool use, const double* xs,
                                               double* y, double* red)
{
    constexpr int R = (kMaxFiberPoints * 3 + kWave - 1) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (!use) {
        for (int i = threadIdx.x; i < bs; i += NW * kWave) y[i] = xs[i];
        return;
    }
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    const int nr = (bs + kWave - 1) / kWave;
    int j = w;
    for (; j + NW < bs; j += 2 * NW) {
        const double x0 = xs[j], x1 = xs[j + NW];
        const T* c0 = B + (size_t)bs * j;
        const T* c1 = B + (size_t)bs * (j + NW);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane + kWave * r;
            if (r < nr && i < bs) acc[r] += (double)c0[i] * x0 + (double)c1[i] * x1;
        }
    }
    for (; j < bs; j += NW) {
        const double x0 = xs[j];
        const T* c0 = B + (size_t)bs * j;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane + kWave * r;
            if (r < nr && i < bs) acc[r] += (double)c0[i] * x0;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + kWave * r;
        if (r < nr && i < bs) red[w * bs + i] = acc[r];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < bs; i += NW * kWave) {
        double v = red[i];
#pragma unroll
        for (int q = 1; q < NW; ++q) v += red[q * bs + i];
        y[i] = v;
    }
}

// The same product in FP32 arithmetic (CYMG_OPT_FP32_APPLY): the operand
// rounded to FP32, one FP32 fused multiply-add per entry, the NW partial sums
// added in FP32 in wave order (red holds them as floats), the result widened.
// Rounding the operand makes the operator nonlinear; the four-launch
// iteration, which applies M^-1 to each vector directly, is the one that uses it.
template <int NW>
__device__ __forceinline__ void block_apply_mw_f32(const float* __restrict__ B, int bs, bool use, const double* xs,
                                                   double* y, double* red)
{
    constexpr int R = (kMaxFiberPoints * 3 + kWave - 1) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (!use) {
        for (int i = threadIdx.x; i < bs; i += NW * kWave) y[i] = xs[i];
        return;
    }
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0f;
    const int nr = (bs + kWave - 1) / kWave;
    for (int j = w; j < bs; j += NW) {
        const float x0 = (float)xs[j];
        const float* c0 = B + (size_t)bs * j;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane + kWave * r;
            if (r < nr && i < bs) acc[r] = fmaf(c0[i], x0, acc[r]);
        }
    }
    float* rf = reinterpret_cast<float*>(red);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + kWave * r;
        if (r < nr && i < bs) rf[w * bs + i] = acc[r];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < bs; i += NW * kWave) {
        float v = rf[i];
#pragma unroll
        for (int q = 1; q < NW; ++q) v += rf[q * bs + i];
        y[i] = (double)v;
    }
}

}  // namespace cymg
