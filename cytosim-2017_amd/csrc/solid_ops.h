// Device-side operators of the Solid Mecable (solid.cc): the rigid-body
// projection of forces (setSpeedsFromForces), and the preconditioner block
// of a Solid kept in low-rank form.  One workgroup handles one Solid; thread t
// owns points t, t + blockDim, ...
//
// A Solid's velocities are rigid motions: with G the map from (t, r) to the
// point velocities t + r x x_p (three columns of each in 3D, two and one in
// 2D), setSpeedsFromForces(X, Y, sc) is Y = G Q(sc) G' X, where G'X is the
// total force and torque about the origin and Q(sc) = sc Q1 is the 6x6 (3x3
// in 2D) reduction through the drag, the centre and the factored momentum
// matrix.  The diagonal block M = I - dt G Q1 G' K_ss (K_ss: the Solid's part
// of mB and mC, meca.cc:231-283) is the identity plus a rank-6 term G A, so
// its inverse is exact in the same form (Sherman-Morrison-Woodbury):
//   M^-1 = I - G W,   W = (I + A G)^-1 A,   A = -dt Q1 (K_ss G)'.
// The Solid keeps W (6 values per unknown) instead of the dense bs x bs
// inverse the reference computes with getrf + getri (meca.cc:352-355): the
// same operator up to rounding, with 6 n instead of n^2 entries to read.
#pragma once

#include "fiber_ops.h"

namespace cymg {

constexpr int kSolidGeo = 16;   // doubles per Solid
enum { SG_DRAG = 0, SG_ROT = 1, SG_MOM2D = 2, SG_C = 3, SG_U = 6, SG_SHAPE_SQR = 12 };
constexpr int kSolidThreads = 256;

struct SolidArgs {
    uint32_t nsol;
    uint32_t p0;            // first Solid point: every fiber point comes before
    uint32_t nfib;          // fibers (the Solid's slot in per-Mecable arrays is nfib + k)
    const uint32_t* np;     // points per Solid
    const uint32_t* idx;    // matIndex per Solid
    const double* rad;      // soRadius per Solid point (point - p0)
    const double* geo;      // kSolidGeo per Solid: soDrag, soDragRot, soMom2D, soCenter, soMom factor U
    const double* W;        // 6 per unknown (unknown - DIM p0): the block inverse's low-rank factor
    const int* useW;        // per Solid: the block could be inverted
    const double* pos;      // the points the projection uses (vPTS = psPos during the solve)
};

template <int D>
struct RigidDim {
    static constexpr int NG = D == 3 ? 6 : 3;   // translations + rotations
};

// row (p, d) of G: the velocity of coordinate d of a point at x under each
// unit translation and unit rotation
template <int D>
__device__ __forceinline__ void rigid_row(const double* x, int d, double g[6])
{
    if constexpr (D == 3) {
        g[0] = d == 0; g[1] = d == 1; g[2] = d == 2;
        // (e_k x x)_d
        g[3] = d == 0 ? 0.0 : (d == 1 ? -x[2] : x[1]);
        g[4] = d == 0 ? x[2] : (d == 1 ? 0.0 : -x[0]);
        g[5] = d == 0 ? -x[1] : (d == 1 ? x[0] : 0.0);
    } else {
        g[0] = d == 0; g[1] = d == 1;
        g[2] = d == 0 ? -x[1] : x[0];
        g[3] = g[4] = g[5] = 0.0;
    }
}

// (T', R') = Q(sc) (T, Rr): the reduction of Solid::setSpeedsFromForces
// (2D: solid.cc:763-782, 3D: 858-888) from the total force T and the torque
// about the origin Rr, with the momentum matrix factor U (dpotrs, 'U')
template <int D>
__device__ __forceinline__ void solid_q(const double* geo, double sc, const double in[6], double out[6])
{
    const double* c = geo + SG_C;
    const double s = sc / geo[SG_DRAG];
    if constexpr (D == 3) {
        const double TX = in[0], TY = in[1], TZ = in[2];
        double R[3] = { sc * (in[3] + (TY * c[2] - TZ * c[1])), sc * (in[4] + (TZ * c[0] - TX * c[2])),
                        sc * (in[5] + (TX * c[1] - TY * c[0])) };
        const double* U = geo + SG_U;   // u00 u01 u02 u11 u12 u22
        // U' y = R
        R[0] = R[0] / U[0];
        R[1] = (R[1] - U[1] * R[0]) / U[3];
        R[2] = (R[2] - U[2] * R[0] - U[4] * R[1]) / U[5];
        // U x = y
        R[2] = R[2] / U[5];
        R[1] = (R[1] - U[4] * R[2]) / U[3];
        R[0] = (R[0] - U[1] * R[1] - U[2] * R[2]) / U[0];
        out[0] = s * TX + (c[1] * R[2] - c[2] * R[1]);
        out[1] = s * TY + (c[2] * R[0] - c[0] * R[2]);
        out[2] = s * TZ + (c[0] * R[1] - c[1] * R[0]);
        out[3] = R[0]; out[4] = R[1]; out[5] = R[2];
    } else {
        const double TX = in[0], TY = in[1];
        const double R = (sc / geo[SG_MOM2D]) * (in[2] + (TX * c[1] - TY * c[0]));
        out[0] = s * TX + c[1] * R;
        out[1] = s * TY - c[0] * R;
        out[2] = R;
        out[3] = out[4] = out[5] = 0.0;
    }
}

// the total force and torque about the origin of forces f at x, accumulated
template <int D>
__device__ __forceinline__ void add_moments(const double* x, const double* f, double acc[6])
{
    if constexpr (D == 3) {
        acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
        acc[3] += x[1] * f[2] - x[2] * f[1];
        acc[4] += x[2] * f[0] - x[0] * f[2];
        acc[5] += x[0] * f[1] - x[1] * f[0];
    } else {
        acc[0] += f[0]; acc[1] += f[1];
        acc[2] += x[0] * f[1] - x[1] * f[0];
    }
}

// velocity of coordinate d of the point at x under the rigid motion q = (t, r)
template <int D>
__device__ __forceinline__ double rigid_velocity(const double* q, const double* x, int d)
{
    if constexpr (D == 3) {
        if (d == 0) return q[0] + q[4] * x[2] - q[5] * x[1];
        if (d == 1) return q[1] + q[5] * x[0] - q[3] * x[2];
        return q[2] + q[3] * x[1] - q[4] * x[0];
    } else {
        return d == 0 ? q[0] - q[2] * x[1] : q[1] + q[2] * x[0];
    }
}

// sums of NV values over the workgroup, in wavefront order (every thread gets
// the totals); red holds NV * 16 doubles
template <int NV>
__device__ __forceinline__ void wg_sum(double v[NV], double* red)
{
    const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if ((threadIdx.x & (kWave - 1)) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[16 * k + w] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int q = 0; q < nw; ++q) s += red[16 * k + q];
        v[k] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ double wg_max(double v, double* red)
{
    const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    v = wave_max(v);
    if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
    __syncthreads();
    double m = red[0];
    for (int q = 1; q < nw; ++q) m = nan_max(m, red[q]);
    __syncthreads();
    return m;
}

// coupling force on point r: [B X]_r (isotropic) + [C X]_r
template <int D, typename Op>
__device__ __forceinline__ void coupling_row(const FiberArgs& a, uint32_t r, const Op& X, double fp[D])
{
#pragma unroll
    for (int d = 0; d < D; ++d) fp[d] = 0.0;
    if (a.Buse) {
        // a point with many links (an aster centre) has a long row: its entries
        // and operands are loaded 8 at a time, then added in order
        const uint32_t e1 = a.Brow[r + 1];
        uint32_t q = a.Brow[r];
        for (; q + 8 <= e1; q += 8) {
            double v[8], x[8][D];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                v[j] = a.Bval[q + j];
                const uint32_t c = a.Bcol[q + j];
#pragma unroll
                for (int d = 0; d < D; ++d) x[j][d] = X(D * c + d);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
#pragma unroll
                for (int d = 0; d < D; ++d) fp[d] += v[j] * x[j][d];
        }
        for (; q < e1; ++q) {
            const double v = a.Bval[q];
            const uint32_t c = a.Bcol[q];
#pragma unroll
            for (int d = 0; d < D; ++d) fp[d] += v * X(D * c + d);
        }
    }
    if (a.Cuse) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double acc = 0.0;
            const uint32_t e1 = a.Crow[D * r + d + 1];
            for (uint32_t q = a.Crow[D * r + d]; q < e1; ++q) acc += a.Cval[q] * X(a.Ccol[q]);
            fp[d] += acc;
        }
    }
}

// last p in [0, n) with rows[p] <= q (rows ascending)
__device__ __forceinline__ uint32_t row_of(const uint32_t* rows, uint32_t n, uint32_t q)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (rows[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// total force and torque of (B + C) X over the Solid's rows.  Only these six
// sums matter to the projection, so the CSR entries of the Solid's rows are
// spread over the workgroup one entry per thread (a Solid point can carry one
// link per fiber of its aster: a row-per-thread loop would serialise them),
// each entry's row found by binary search in the row extents staged in LDS
// (lrow: cap entries).  Solids too large for the staging use one row per thread.
template <int D, typename Op>
__device__ __forceinline__ void solid_coupling_moments(const FiberArgs& a, const SolidArgs& s, uint32_t k,
                                                       const Op& X, double acc[6], uint32_t* lrow, uint32_t cap)
{
    const uint32_t np = s.np[k], idx = s.idx[k];
    if (a.Buse) {
        if (np + 1 <= cap) {
            for (uint32_t i = threadIdx.x; i <= np; i += blockDim.x) lrow[i] = a.Brow[idx + i];
            __syncthreads();
            // each thread a contiguous run of entries: one binary search for
            // its first row, then the rows followed forward; the entries and
            // their operands loaded four at a time (addresses clamped to the
            // run's last entry, weights beyond it zero) before they are used
            const uint32_t q0 = lrow[0], q1 = lrow[np];
            const uint32_t per = (q1 - q0 + blockDim.x - 1) / blockDim.x;
            const uint32_t qb = q0 + threadIdx.x * per;
            const uint32_t qe = qb + per < q1 ? qb + per : q1;
            if (qb < qe) {
                uint32_t rr = row_of(lrow, np, qb);
                const uint32_t ql = qe - 1;
                for (uint32_t q = qb; q < qe; q += 4) {
                    double v[4], x[4][D];
                    uint32_t c[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t qq = q + j < ql ? q + j : ql;
                        const double w = a.Bval[qq];
                        v[j] = q + j < qe ? w : 0.0;
                        c[j] = a.Bcol[qq];
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j)
#pragma unroll
                        for (int d = 0; d < D; ++d) x[j][d] = X(D * c[j] + d);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (q + j >= qe) break;
                        while (lrow[rr + 1] <= q + j) ++rr;
                        double f[D];
#pragma unroll
                        for (int d = 0; d < D; ++d) f[d] = v[j] * x[j][d];
                        add_moments<D>(s.pos + D * (idx + rr), f, acc);
                    }
                }
            }
            __syncthreads();
        } else {
            for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
                double f[D];
#pragma unroll
                for (int d = 0; d < D; ++d) f[d] = 0.0;
                const uint32_t r = idx + p;
                for (uint32_t q = a.Brow[r]; q < a.Brow[r + 1]; ++q) {
                    const double v = a.Bval[q];
                    const uint32_t c = a.Bcol[q];
#pragma unroll
                    for (int d = 0; d < D; ++d) f[d] += v * X(D * c + d);
                }
                add_moments<D>(s.pos + D * r, f, acc);
            }
        }
    }
    if (a.Cuse) {
        const uint32_t nd = D * np, o = D * idx;
        if (nd + 1 <= cap) {
            for (uint32_t i = threadIdx.x; i <= nd; i += blockDim.x) lrow[i] = a.Crow[o + i];
            __syncthreads();
            const uint32_t q1 = lrow[nd];
            for (uint32_t q = lrow[0] + threadIdx.x; q < q1; q += blockDim.x) {
                const uint32_t i = row_of(lrow, nd, q);
                const uint32_t p = i / D;
                double f[D];
#pragma unroll
                for (int d = 0; d < D; ++d) f[d] = 0.0;
                f[i - D * p] = a.Cval[q] * X(a.Ccol[q]);
                add_moments<D>(s.pos + D * (idx + p), f, acc);
            }
            __syncthreads();
        } else {
            for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
                double f[D];
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    double t = 0.0;
                    for (uint32_t q = a.Crow[o + D * p + d]; q < a.Crow[o + D * p + d + 1]; ++q)
                        t += a.Cval[q] * X(a.Ccol[q]);
                    f[d] = t;
                }
                add_moments<D>(s.pos + D * (idx + p), f, acc);
            }
        }
    }
}

// Meca::multiply on the rows of Solid k: Y = X + setSpeedsFromForces((B+C) X, -dt)
// (meca.cc:157-181); out(i, value) receives each row.  lrow: LDS staging of
// cap row extents (see solid_coupling_moments).
template <int D, typename Op, typename Out>
__device__ __forceinline__ void solid_operator(const FiberArgs& a, const SolidArgs& s, uint32_t k, const Op& X,
                                               const Out& out, double* red, uint32_t* lrow, uint32_t cap)
{
    constexpr int NG = RigidDim<D>::NG;
    const uint32_t np = s.np[k], idx = s.idx[k];
    const double* geo = s.geo + kSolidGeo * k;
    double acc[6] = { 0, 0, 0, 0, 0, 0 };
    solid_coupling_moments<D>(a, s, k, X, acc, lrow, cap);
    wg_sum<NG>(acc, red);
    double q[6];
    solid_q<D>(geo, -a.dt, acc, q);
    for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
        const double* x = s.pos + D * (idx + p);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const uint32_t i = D * (idx + p) + d;
            out(i, X(i) + rigid_velocity<D>(q, x, d));
        }
    }
}

// y = M^-1 x on the rows of Solid k (a copy when the block was not
// invertible, meca.cc:463); xin(i) reads the operand, out(i, v) writes
template <int D, typename In, typename Out>
__device__ __forceinline__ void solid_precond(const SolidArgs& s, uint32_t k, const In& xin, const Out& out,
                                              double* red)
{
    constexpr int NG = RigidDim<D>::NG;
    const uint32_t np = s.np[k], idx = s.idx[k];
    const uint32_t o = D * idx, n = D * np;
    if (!s.useW[k]) {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) out(o + i, xin(o + i));
        return;
    }
    const double* W = s.W + (size_t)6 * (o - D * s.p0);
    double acc[6] = { 0, 0, 0, 0, 0, 0 };
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double xi = xin(o + i);
#pragma unroll
        for (int c = 0; c < NG; ++c) acc[c] += W[6 * i + c] * xi;
    }
    wg_sum<NG>(acc, red);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t p = i / D;
        const int d = (int)(i - D * p);
        out(o + i, xin(o + i) - rigid_velocity<D>(acc, s.pos + D * (idx + p), d));
    }
}

}  // namespace cymg
