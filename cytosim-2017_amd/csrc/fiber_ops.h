// Device-side per-fiber operators of the implicit step.  One workgroup of one
// wavefront handles one fiber; lane p owns point p (p, p+64 for fibers longer
// than a wavefront).  The fiber's coordinates, forces and Lagrange multipliers
// are staged in LDS; the coupling rows of mB/mC are gathered from HBM.
#pragma once

#include "common.h"

namespace cymg {

// Everything a per-fiber kernel needs, passed by value.
// Levels of parallel cyclic reduction that cover a tridiagonal system of up
// to kWave rows (steps 1, 2, 4, ..., 32).
constexpr int kPcrLevels = 6;

struct FiberArgs {
    const uint32_t* np;       // points per fiber
    const uint32_t* idx;      // matIndex
    const uint32_t* seg;      // first segment
    const uint32_t* fib_sys;  // system of the fiber
    const uint32_t* sys_begin;// first fiber of each system (nsys+1)
    const double* mob;        // rfMobility
    const double* rfrig;      // rfRigidity = rigidity / fnCut^3
    const double* cut;        // fnCut
    const double* diff;       // rfDiff, DIM per segment
    const double* jd;         // pttrf diagonal of JJ'
    const double* je;         // pttrf subdiagonal of L
    const double* jd0;        // JJ' diagonal 2|d_s|^2 (unfactored)
    const double* je0;        // JJ' off-diagonal -d_s.d_{s+1} (unfactored)
    const double* pdiff;      // mtJJtiJforce (P' coefficients)
    const double* pcr;        // PCR coefficients of JJ' (fibers of at most kWave segments):
                              // [2l][seg] alpha and [2l+1][seg] gamma of level l < kPcrLevels,
                              // [2 kPcrLevels][seg] the reduced diagonal
    const uint32_t* pfib;     // fiber of each point
    uint32_t nseg;            // segments in total
    // full symmetric CSR of mB (rows = points) and mC (rows = dofs)
    const uint32_t* Brow; const uint32_t* Bcol; const double* Bval; int Buse;
    const uint32_t* Crow; const uint32_t* Ccol; const double* Cval; int Cuse;
    double dt;
};

// P: the longest fiber the workgroup takes (kShortFiberPoints for the
// launches of layouts without longer fibers: less LDS, more workgroups per CU)
template <int D, int P = kMaxFiberPoints>
struct FiberLds {
    double x[P * D];     // fiber coordinates of the operand
    double f[P * D];     // forces
    double y[P * D];     // result
    double t[P];         // J*f, then multipliers
    double d[P];         // pttrf D
    double e[P];         // pttrf L subdiagonal
    double red[kWave];
    uint32_t hub[P];     // points whose mB row is long (many links on one point)
    uint32_t nhub;
};

// mB rows longer than this are gathered by the whole workgroup, one entry per
// thread, instead of by the thread that owns the point
constexpr uint32_t kHubRow = 24;

// Wavefront reductions (every lane of the wavefront must be active): DPP
// moves inside each row of 16 lanes (the lane xor 1 and xor 2 quad
// permutations, then the half-row mirror i <-> 7 - i and the row mirror
// i <-> 15 - i), after which every lane of a row holds the row's result; the
// four row results are read with readlane and combined in row order.  Every
// lane returns the same value.  A butterfly of cross-lane shuffles took six
// LDS-crossbar round trips; these moves are register operations.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_i32<CTRL>((int)(unsigned)(b & 0xffffffffull));
    const unsigned hi = (unsigned)dpp_i32<CTRL>((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppRowMirror = 0x140;
constexpr int kDppHalfMirror = 0x141;

template <typename Op>
__device__ __forceinline__ double wave_reduce(double v, const Op& op)
{
    v = op(v, dpp_f64<kDppXor1>(v));
    v = op(v, dpp_f64<kDppXor2>(v));
    v = op(v, dpp_f64<kDppHalfMirror>(v));
    v = op(v, dpp_f64<kDppRowMirror>(v));
    const double r0 = readlane_f64(v, 0), r1 = readlane_f64(v, 16);
    const double r2 = readlane_f64(v, 32), r3 = readlane_f64(v, 48);
    return op(op(r0, r1), op(r2, r3));
}

template <typename Op>
__device__ __forceinline__ int wave_reduce_i32(int v, const Op& op)
{
    v = op(v, dpp_i32<kDppXor1>(v));
    v = op(v, dpp_i32<kDppXor2>(v));
    v = op(v, dpp_i32<kDppHalfMirror>(v));
    v = op(v, dpp_i32<kDppRowMirror>(v));
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return op(op(r0, r1), op(r2, r3));
}

__device__ __forceinline__ double wave_sum(double v)
{
    return wave_reduce(v, [](double a, double b) { return a + b; });
}

__device__ __forceinline__ double nan_max(double a, double b)
{
    // maximum that propagates NaN
    if (a != a || b != b) return __builtin_nan("");
    return a > b ? a : b;
}

__device__ __forceinline__ double wave_max(double v)
{
    return wave_reduce(v, [](double a, double b) { return nan_max(a, b); });
}

__device__ __forceinline__ double wave_min(double v)
{
    return wave_reduce(v, [](double a, double b) { return fmin(a, b); });
}

__device__ __forceinline__ int wave_max_i32(int v)
{
    return wave_reduce_i32(v, [](int a, int b) { return a > b ? a : b; });
}

// dptts2 with one right-hand side (LAPACK): solve L*D*L' x = b, sequential
__device__ __forceinline__ void ptts2_serial(int n, const double* d, const double* e, double* b)
{
    if (n <= 1) {
        if (n == 1) b[0] *= 1.0 / d[0];
        return;
    }
    double prev = b[0];
    for (int i = 1; i < n; ++i) {
        prev = b[i] - prev * e[i - 1];
        b[i] = prev;
    }
    double nxt = b[n - 1] / d[n - 1];
    b[n - 1] = nxt;
    for (int i = n - 2; i >= 0; --i) {
        nxt = b[i] / d[i] - nxt * e[i];
        b[i] = nxt;
    }
}

// Parallel cyclic reduction of the symmetric tridiagonal system JJ' x = t
// across one wavefront (lane = row, S <= 64): log2(S) steps of 8 shuffles.
// The same system as dptts2 solves with the pttrf factors; rounding differs
// at the ulp level.  Lanes >= S hold identity rows.  With span < S the system
// must be block diagonal with blocks of at most span rows (zero couplings
// between blocks): span bounds the number of steps.
__device__ __forceinline__ double pcr_solve_span(int S, int span, double a, double b, double c, double d)
{
    const int lane = threadIdx.x & (kWave - 1);
    if (lane >= S) { a = 0.0; b = 1.0; c = 0.0; d = 0.0; }
    for (int k = 1; k < span; k <<= 1) {
        double al = __shfl_up(a, k, kWave), bl = __shfl_up(b, k, kWave);
        double cl = __shfl_up(c, k, kWave), dl = __shfl_up(d, k, kWave);
        double au = __shfl_down(a, k, kWave), bu = __shfl_down(b, k, kWave);
        double cu = __shfl_down(c, k, kWave), du = __shfl_down(d, k, kWave);
        if (lane < k) { al = 0.0; bl = 1.0; cl = 0.0; dl = 0.0; }
        if (lane + k >= kWave) { au = 0.0; bu = 1.0; cu = 0.0; du = 0.0; }
        const double alpha = -a / bl;
        const double gamma = -c / bu;
        b = b + alpha * cl + gamma * au;
        d = d + alpha * dl + gamma * du;
        a = alpha * al;
        c = gamma * cu;
    }
    return d / b;
}

__device__ __forceinline__ double pcr_solve(int S, double a, double b, double c, double d)
{
    return pcr_solve_span(S, S, a, b, c, d);
}

// Load the fiber's pttrf factors into LDS.
// (only the serial path for fibers longer than a wavefront needs them, unless forced)
template <int D, int P>
__device__ __forceinline__ void load_factors(const FiberArgs& a, FiberLds<D, P>& s, uint32_t seg, int S,
                                             bool force = false)
{
    if (!force && S <= kWave) return;
    for (int k = threadIdx.x; k < S; k += blockDim.x) {
        s.d[k] = a.jd[seg + k];
        s.e[k] = (k + 1 < S) ? a.je[seg + k] : 0.0;
    }
}

// f_p = [R X]_p + [B X]_p + [C X]_p + [P' X]_p for the points of the fiber.
// X: global operand (coupling gather); s.x must hold the fiber's slice of X.
// operand read through a plain array
struct Plain {
    const double* __restrict__ X;
    __device__ __forceinline__ double operator()(uint32_t j) const { return X[j]; }
};

template <int D, int P, typename Op>
__device__ __forceinline__ void fiber_forces_op(const FiberArgs& a, FiberLds<D, P>& s, const Op& X,
                                                uint32_t np, uint32_t idx, uint32_t seg, double k,
                                                bool rigidity, bool with_pdiff)
{
    const int S = (int)np - 1;
    if (threadIdx.x == 0) s.nhub = 0;
    __syncthreads();
    for (int p = threadIdx.x; p < (int)np; p += blockDim.x) {
        double fp[D];
#pragma unroll
        for (int d = 0; d < D; ++d) fp[d] = 0.0;
        const uint32_t r = idx + p;
        uint32_t cb[D + 1];
        if (a.Cuse) {
#pragma unroll
            for (int d = 0; d <= D; ++d) cb[d] = a.Crow[D * r + d];
        }
        if (a.Buse) {
            uint32_t e1 = a.Brow[r + 1];
            const uint32_t e0 = a.Brow[r];
            if (e1 - e0 > kHubRow) {
                s.hub[atomicAdd(&s.nhub, 1u)] = (uint32_t)p;
                e1 = e0;
            }
            for (uint32_t q = e0; q < e1; ++q) {
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
                for (uint32_t q = cb[d]; q < cb[d + 1]; ++q) acc += a.Cval[q] * X(a.Ccol[q]);
                fp[d] += acc;
            }
        }
        if (rigidity && np > 2) {
            // bending stencil v_j = k (x_j - 2 x_{j+1} + x_{j+2}); Y_j -= v, Y_{j+1} += 2v, Y_{j+2} -= v
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double acc = 0.0;
                if (p <= (int)np - 3) {
                    const double v = k * (s.x[D * p + d] - 2 * s.x[D * (p + 1) + d] + s.x[D * (p + 2) + d]);
                    acc -= v;
                }
                if (p >= 1 && p <= (int)np - 2) {
                    const double v = k * (s.x[D * (p - 1) + d] - 2 * s.x[D * p + d] + s.x[D * (p + 1) + d]);
                    acc += v + v;
                }
                if (p >= 2) {
                    const double v = k * (s.x[D * (p - 2) + d] - 2 * s.x[D * (p - 1) + d] + s.x[D * p + d]);
                    acc -= v;
                }
                fp[d] += acc;
            }
        }
        if (with_pdiff) {
            // P': w = c_s (X_{s+1} - X_s); Y_s += w; Y_{s+1} -= w
            if (p < S) {
                const double c = a.pdiff[seg + p];
                if (c != 0.0)
#pragma unroll
                    for (int d = 0; d < D; ++d) fp[d] += c * (s.x[D * (p + 1) + d] - s.x[D * p + d]);
            }
            if (p >= 1) {
                const double c = a.pdiff[seg + p - 1];
                if (c != 0.0)
#pragma unroll
                    for (int d = 0; d < D; ++d) fp[d] -= c * (s.x[D * p + d] - s.x[D * (p - 1) + d]);
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d) s.f[D * p + d] = fp[d];
    }
    __syncthreads();
    // long rows: entries spread over the workgroup, summed in a fixed order
    const uint32_t nh = s.nhub;
    for (uint32_t h = 0; h < nh; ++h) {
        const uint32_t p = s.hub[h], r = idx + p;
        double acc[D];
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d] = 0.0;
        const uint32_t e1 = a.Brow[r + 1];
        // four entries per thread per round: their loads are issued together
        constexpr uint32_t U = 4;
        const uint32_t stride = blockDim.x;
        const uint32_t el = e1 - 1;
        for (uint32_t q0 = a.Brow[r] + threadIdx.x; q0 < e1; q0 += U * stride) {
            double v[U];
            uint32_t c[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t q = q0 + u * stride;
                const uint32_t qc = q < el ? q : el;
                v[u] = q < e1 ? a.Bval[qc] : 0.0;
                c[u] = a.Bcol[qc];
            }
            double xv[U][D];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
#pragma unroll
                for (int d = 0; d < D; ++d) xv[u][d] = X(D * c[u] + d);
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
#pragma unroll
                for (int d = 0; d < D; ++d) acc[d] += v[u] * xv[u][d];
        }
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d] = wave_sum(acc[d]);
        const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
        if ((threadIdx.x & (kWave - 1)) == 0)
#pragma unroll
            for (int d = 0; d < D; ++d) s.red[D * w + d] = acc[d];
        __syncthreads();
        if (threadIdx.x == 0)
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double t = 0.0;
                for (int q = 0; q < nw; ++q) t += s.red[D * q + d];
                s.f[D * p + d] += t;
            }
        __syncthreads();
    }
}

template <int D, int P>
__device__ __forceinline__ void fiber_forces(const FiberArgs& a, FiberLds<D, P>& s, const double* __restrict__ X,
                                             uint32_t np, uint32_t idx, uint32_t seg, double k,
                                             bool rigidity, bool with_pdiff)
{
    fiber_forces_op<D>(a, s, Plain{ X }, np, idx, seg, k, rigidity, with_pdiff);
}

// s.y = sca * P s.f   with P = I - J'(JJ')^-1 J; the multipliers (JJ')^-1 J f
// are left in s.t (rigid_fiber_project.cc:95-149).  s.d/s.e hold the factors.
template <int D, int P>
__device__ __forceinline__ void fiber_project(const FiberArgs& a, FiberLds<D, P>& s, uint32_t np, uint32_t seg,
                                              double sca)
{
    const int S = (int)np - 1;
    if (S <= kWave) {
        const int q = threadIdx.x;
        double t = 0.0, sa = 0.0, sb = 1.0, sc = 0.0;
        if (q < S) {
#pragma unroll
            for (int d = 0; d < D; ++d) t += a.diff[D * (seg + q) + d] * (s.f[D * (q + 1) + d] - s.f[D * q + d]);
            sb = a.jd0[seg + q];
            if (q > 0) sa = a.je0[seg + q - 1];
            if (q + 1 < S) sc = a.je0[seg + q];
        }
        const double lam = pcr_solve(S, sa, sb, sc, t);
        if (q < S) s.t[q] = lam;
    } else {
        for (int q = threadIdx.x; q < S; q += blockDim.x) {
            double t = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) t += a.diff[D * (seg + q) + d] * (s.f[D * (q + 1) + d] - s.f[D * q + d]);
            s.t[q] = t;
        }
        __syncthreads();
        if (threadIdx.x == 0) ptts2_serial(S, s.d, s.e, s.t);
    }
    __syncthreads();
    for (int p = threadIdx.x; p < (int)np; p += blockDim.x) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double v = s.f[D * p + d];
            if (p < S) v = v + a.diff[D * (seg + p) + d] * s.t[p];
            if (p >= 1) v = v - a.diff[D * (seg + p - 1) + d] * s.t[p - 1];
            s.y[D * p + d] = sca * v;
        }
    }
}

// s.y = X - dt*mob*P (R + B + C + P') X  (Meca::multiply, meca.cc:157-181),
// for one fiber; s.x must hold the fiber's slice of X, X(j) gives any entry.
template <int D, int P, typename Op>
__device__ __forceinline__ void fiber_operator_op(const FiberArgs& a, FiberLds<D, P>& s, const Op& X, uint32_t f)
{
    const uint32_t np = a.np[f], idx = a.idx[f], seg = a.seg[f];
    fiber_forces_op<D>(a, s, X, np, idx, seg, a.rfrig[f], true, true);
    __syncthreads();
    fiber_project<D>(a, s, np, seg, -a.dt * a.mob[f]);
    __syncthreads();
    for (int i = threadIdx.x; i < (int)(D * np); i += blockDim.x) s.y[i] = s.y[i] + s.x[i];
    __syncthreads();
}

template <int D, int P>
__device__ __forceinline__ void fiber_operator(const FiberArgs& a, FiberLds<D, P>& s, const double* __restrict__ X,
                                               uint32_t f)
{
    fiber_operator_op<D>(a, s, Plain{ X }, f);
}

}  // namespace cymg
