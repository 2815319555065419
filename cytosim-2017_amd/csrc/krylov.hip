// Preconditioned BiCGStab of the implicit step (Solver::BCGSP,
// bicgstab.h:247-338; Solver::BCGS, bicgstab.h:163-243; Monitor,
// bicgstab.h:51-92 and bicgstab.cc:9-49), entirely on the device.
//
// After an init launch, one iteration is four fused per-fiber launches:
//   A: omega, next rho, beta from D's partials; [x += w shat; r = s - w t];
//      |r|_inf partials; p = r + beta (p - w v); phat = Binv p
//   B: convergence test on A's partials; v = M phat; <rt, v> partials
//   C: alpha = rho / <rt, v>; s = r - alpha v; x += alpha phat; |s|_inf
//      partials; shat = Binv s
//   D: t = M shat; <t,t>, <t,s>, <rt,s>, <rt,t> partials
// The next rho is <rt, r_new> = <rt, s> - omega <rt, t>, so no extra pass over
// r is needed.  No kernel synchronises with other workgroups: each workgroup
// writes the partial sums of its fiber, and every workgroup of the next launch
// re-reduces the partials of its system in fiber order, so all of them derive
// bit-identical scalars without atomics or device-scope fences (kernel
// boundaries order the memory).  The first workgroup of each system writes the
// scalar state into the other half of a double buffer.  The host reads the
// stop flag between chunks of iterations.
#include <algorithm>

#include "args.h"
#include "pc_ops.h"
#include "tile_ops.h"

namespace cymg {

constexpr int kApplyWaves = 4;
constexpr int kApplyThreads = kApplyWaves * kWave;

enum { ST_INIT = 0, ST_A = 1, ST_B = 2, ST_C = 3, ST_D = 4 };

struct KArgs {
    FiberArgs fa;
    const uint64_t* blkoff;
    const int* useblk;
    const double* B64;
    const float* B32;
    const double* b;
    double* x;
    double* r;
    double* rt;
    double* p;
    double* v;
    double* t;
    double* ph;
    double* sh;
    // fused variant
    double* y;          // Binv r
    double* w;          // Binv t
    double* z_in;       // Binv v of the previous iteration
    double* z_out;
    double* ph_in;      // phat of the previous iteration
    double* ph_out;
    double* part;       // partial sums [slot][fiber] (four launches) or [slot][tile] (two)
    double* part4;      // [fiber or tile][4] partials of D
    const double* fpart;// per-fiber partials of the setup (noise, |b|_inf)
    TileArgs ta;
    SysScalars* st_in;
    SysScalars* st_out;
    uint32_t nfib;
    uint32_t nsys;
};

// Reductions over the partials of one system (entries [fb, fe)), repeated by
// every workgroup in the same order, so all of them derive the same value.
// Each lane issues up to kRed loads before using any (one memory round trip
// for up to kRed * 64 entries), then the wavefront reduces with xor shuffles.
constexpr int kRed = 8;

__device__ __forceinline__ double wsum(const double* part, uint32_t fb, uint32_t fe)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    double v = 0.0;
    for (uint32_t base = fb; base < fe; base += kRed * kWave) {
        double t[kRed];
#pragma unroll
        for (int u = 0; u < kRed; ++u) {
            const uint32_t i = base + u * kWave + lane;
            t[u] = i < fe ? part[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kRed; ++u) v += t[u];
    }
    return wave_sum(v);
}

__device__ __forceinline__ double wmax(const double* part, uint32_t fb, uint32_t fe)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    double v = 0.0;
    for (uint32_t base = fb; base < fe; base += kRed * kWave) {
        double t[kRed];
#pragma unroll
        for (int u = 0; u < kRed; ++u) {
            const uint32_t i = base + u * kWave + lane;
            t[u] = i < fe ? part[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kRed; ++u) v = nan_max(v, t[u]);
    }
    return wave_max(v);
}

__device__ __forceinline__ double wmin(const double* part, uint32_t fb, uint32_t fe)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    double v = __builtin_inf();
    for (uint32_t base = fb; base < fe; base += kRed * kWave) {
        double t[kRed];
#pragma unroll
        for (int u = 0; u < kRed; ++u) {
            const uint32_t i = base + u * kWave + lane;
            t[u] = i < fe ? part[i] : __builtin_inf();
        }
#pragma unroll
        for (int u = 0; u < kRed; ++u) v = fmin(v, t[u]);
    }
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}

// the 4 partials of D, stored together per entry ([entry][4])
__device__ __forceinline__ void wsum4(const double* part4, uint32_t fb, uint32_t fe, double out[4])
{
    constexpr int U = kRed / 2;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    double v[4] = { 0, 0, 0, 0 };
    for (uint32_t base = fb; base < fe; base += U * kWave) {
        double4 t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = base + u * kWave + lane;
            t[u] = i < fe ? *reinterpret_cast<const double4*>(part4 + 4 * (size_t)i) : make_double4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            v[0] += t[u].x; v[1] += t[u].y; v[2] += t[u].z; v[3] += t[u].w;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q] = wave_sum(v[q]);
}

// Monitor::finished on |r|_inf (bicgstab.cc:38-48), then the rho = 0 breakdown
__device__ __forceinline__ void decide(SysScalars& s, double res)
{
    if (res != res) {
        s.resid = __builtin_inf();
        s.done = 1;
        return;
    }
    s.resid = res;
    if (s.iter > s.iter_max) s.done = 1;
    else if (res < s.tol) s.done = 1;
    else if (s.rho1 == 0.0) { s.done = 1; s.flag = 2; }
}

__device__ __forceinline__ bool first_of_system(const FiberArgs& a, uint32_t sys)
{
    return blockIdx.x == a.sys_begin[sys];
}

template <int PC>
__device__ __forceinline__ void apply_block(const KArgs& k, uint32_t f, int bs, const double* xs, double* y,
                                            double* red)
{
    if constexpr (PC == CYMG_PRECOND_FP64)
        block_apply_mw<double, kApplyWaves>(k.B64 + k.blkoff[f], bs, k.useblk[f] != 0, xs, y, red);
    else if constexpr (PC == CYMG_PRECOND_FP32)
        block_apply_mw<float, kApplyWaves>(k.B32 + k.blkoff[f], bs, k.useblk[f] != 0, xs, y, red);
}

// r = b - M x0 (or b), rt = r, x = x0 (or 0); <rt, r> partials; threshold
template <int D, int PC>
__global__ __launch_bounds__(kWave) void k_init(KArgs k, const double* __restrict__ x0, double tolerance,
                                                int tol_mode, int iter_max, double dt)
{
    __shared__ FiberLds<D> s;
    const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int n = D * (int)np;
    if (x0) {
        for (int i = threadIdx.x; i < n; i += kWave) s.x[i] = x0[o + i];
        load_factors<D>(a, s, a.seg[f], (int)np - 1);
        __syncthreads();
        fiber_operator<D>(a, s, x0, f);
    }
    double rr = 0.0;
    for (int i = threadIdx.x; i < n; i += kWave) {
        double ri = k.b[o + i];
        if (x0) ri = ri + (-1.0) * s.y[i];
        k.r[o + i] = ri;
        k.rt[o + i] = ri;
        k.x[o + i] = x0 ? s.x[i] : 0.0;
        rr += ri * ri;
    }
    rr = wave_sum(rr);
    if (threadIdx.x == 0) k.part[P_RR * k.nfib + f] = rr;
    const uint32_t sys = k.nsys == 1 ? 0u : a.fib_sys[f];
    if (first_of_system(a, sys)) {
        const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
        const double noise = wmin(k.part + P_NOISE * k.nfib, fb, fe);
        const double bn = wmax(k.part + P_BNORM * k.nfib, fb, fe);
        if (threadIdx.x == 0) {
            SysScalars w{};
            w.noise = noise;
            w.bnorm = bn;
            w.rho1 = 1.0;
            w.alpha = 0.0;
            w.omega = 1.0;
            w.tol = (tol_mode == CYMG_TOL_RELATIVE) ? tolerance * bn : tolerance * (noise * dt);
            w.resid = __builtin_inf();
            w.resid_s = __builtin_inf();
            w.iter = 0;
            w.iter_max = iter_max;
            w.stage = ST_INIT;
            k.st_out[sys] = w;
        }
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kApplyThreads) void k_iter_a(KArgs k)
{
    constexpr int R = (kMaxFiberPoints * D + kApplyThreads - 1) / kApplyThreads;
    __shared__ double xs[kMaxFiberPoints * D];
    __shared__ double red[kApplyWaves * kMaxFiberPoints * D];
    __shared__ double wr[kApplyWaves];
    __shared__ double bc[4];
    const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const uint32_t sys = k.nsys == 1 ? 0u : a.fib_sys[f];
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int n = D * (int)np;
    // the fiber's vectors are loaded while the scalar state is derived
    double vr[R], vt[R], vs[R], vp[R], vv[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int i = threadIdx.x + q * kApplyThreads;
        if (i < n) {
            vr[q] = k.r[o + i];
            vt[q] = k.t[o + i];
            vs[q] = PC ? k.sh[o + i] : vr[q];
            vp[q] = k.p[o + i];
            vv[q] = k.v[o + i];
        }
    }
    SysScalars sc = k.st_in[sys];
    const bool first = first_of_system(a, sys);
    if (sc.done) {
        if (first && threadIdx.x == 0) k.st_out[sys] = sc;
        return;
    }
    const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
    sc.pending = 0;
    if (threadIdx.x < kWave) {
        if (sc.stage == ST_INIT) {
            const double rr = wsum(k.part + P_RR * k.nfib, fb, fe);
            if (threadIdx.x == 0) bc[0] = rr;
        } else {
            double S4[4];
            wsum4(k.part4, fb, fe, S4);
            if (threadIdx.x == 0) { bc[0] = S4[0]; bc[1] = S4[1]; bc[2] = S4[2]; bc[3] = S4[3]; }
        }
    }
    __syncthreads();
    if (sc.stage == ST_INIT) {
        // first iteration: rho2 = 1, rho1 = <rt, r0>
        sc.rho2 = sc.rho1;
        sc.rho1 = bc[0];
        sc.beta = (sc.rho1 / sc.rho2) * (sc.alpha / sc.omega);
    } else {
        const double tt = bc[0], ts = bc[1], rts = bc[2], rtt = bc[3];
        if (tt == 0.0) {
            sc.done = 1; sc.flag = 0; sc.resid = sc.resid_s;
        } else {
            const double om = ts / tt;
            if (om == 0.0) {
                sc.done = 1; sc.flag = 3; sc.resid = sc.resid_s;
            } else {
                sc.omega = om;
                sc.pending = 1;
                sc.rho2 = sc.rho1;
                sc.rho1 = rts - om * rtt;
                sc.beta = (sc.rho1 / sc.rho2) * (sc.alpha / sc.omega);
                sc.iter += 1;
            }
        }
    }
    sc.stage = ST_A;
    if (first && threadIdx.x == 0) k.st_out[sys] = sc;
    if (sc.done) return;
    double rmax = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int i = threadIdx.x + q * kApplyThreads;
        if (i >= n) continue;
        double ri = vr[q];
        if (sc.pending) {
            // x += omega shat (BCGS: shat = s), r = s - omega t
            k.x[o + i] = k.x[o + i] + sc.omega * vs[q];
            ri = ri + (-sc.omega) * vt[q];
            k.r[o + i] = ri;
        }
        rmax = nan_max(rmax, fabs(ri));
        double pi;
        if (sc.beta == 0.0) pi = ri;
        else pi = sc.beta * (vp[q] + (-sc.omega) * vv[q]) + ri;
        k.p[o + i] = pi;
        xs[i] = pi;
    }
    rmax = wave_max(rmax);
    if ((threadIdx.x & (kWave - 1)) == 0) wr[threadIdx.x / kWave] = rmax;
    __syncthreads();
    if constexpr (PC != CYMG_PRECOND_NONE) apply_block<PC>(k, f, n, xs, k.ph + o, red);
    if (threadIdx.x == 0) {
        double m = wr[0];
        for (int q = 1; q < kApplyWaves; ++q) m = nan_max(m, wr[q]);
        k.part[P_RMAX * k.nfib + f] = m;
    }
}

// convergence test of the last A, one workgroup per system, in place
__global__ __launch_bounds__(kWave) void k_decide(KArgs k, uint32_t nsys)
{
    const uint32_t sys = blockIdx.x;
    if (sys >= nsys) return;
    SysScalars sc = k.st_in[sys];
    if (sc.done || sc.stage != ST_A) return;
    const uint32_t fb = k.fa.sys_begin[sys], fe = k.fa.sys_begin[sys + 1];
    const double res = wmax(k.part + P_RMAX * k.nfib, fb, fe);
    if (threadIdx.x == 0) {
        decide(sc, res);
        k.st_in[sys] = sc;
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kWave) void k_iter_b(KArgs k)
{
    __shared__ FiberLds<D> s;
        const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const uint32_t sys = k.nsys == 1 ? 0u : a.fib_sys[f];
    // the operand slice is loaded while the scalar state is derived
    const double* X = PC ? k.ph : k.p;
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int n = D * (int)np;
    for (int i = threadIdx.x; i < n; i += kWave) s.x[i] = X[o + i];
    SysScalars sc = k.st_in[sys];
    const bool first = first_of_system(a, sys);
    if (!sc.done) {
        const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
        decide(sc, wmax(k.part + P_RMAX * k.nfib, fb, fe));
        sc.stage = ST_B;
    }
    if (first && threadIdx.x == 0) k.st_out[sys] = sc;
    if (sc.done) return;
    load_factors<D>(a, s, a.seg[f], (int)np - 1);
    __syncthreads();
    fiber_operator<D>(a, s, X, f);
    double dl = 0.0;
    for (int i = threadIdx.x; i < n; i += kWave) {
        k.v[o + i] = s.y[i];
        dl += k.rt[o + i] * s.y[i];
    }
    dl = wave_sum(dl);
    if (threadIdx.x == 0) k.part[P_DELTA * k.nfib + f] = dl;
}

template <int D, int PC>
__global__ __launch_bounds__(kApplyThreads) void k_iter_c(KArgs k)
{
    constexpr int R = (kMaxFiberPoints * D + kApplyThreads - 1) / kApplyThreads;
    __shared__ double xs[kMaxFiberPoints * D];
    __shared__ double red[kApplyWaves * kMaxFiberPoints * D];
    __shared__ double wr[kApplyWaves];
    __shared__ double bc[4];
    const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const uint32_t sys = k.nsys == 1 ? 0u : a.fib_sys[f];
    const double* X = PC ? k.ph : k.p;
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int n = D * (int)np;
    double vr[R], vv[R], vx[R], vX[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int i = threadIdx.x + q * kApplyThreads;
        if (i < n) {
            vr[q] = k.r[o + i];
            vv[q] = k.v[o + i];
            vx[q] = k.x[o + i];
            vX[q] = X[o + i];
        }
    }
    SysScalars sc = k.st_in[sys];
    const bool first = first_of_system(a, sys);
    if (!sc.done) {
        const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
        if (threadIdx.x < kWave) {
            const double dsum = wsum(k.part + P_DELTA * k.nfib, fb, fe);
            if (threadIdx.x == 0) bc[0] = dsum;
        }
        __syncthreads();
        const double delta = bc[0];
        if (PC != CYMG_PRECOND_NONE && delta == 0.0) { sc.done = 1; sc.flag = 4; }
        else sc.alpha = sc.rho1 / delta;
        sc.stage = ST_C;
    }
    if (first && threadIdx.x == 0) k.st_out[sys] = sc;
    if (sc.done) return;
    double smax = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int i = threadIdx.x + q * kApplyThreads;
        if (i >= n) continue;
        const double si = vr[q] + (-sc.alpha) * vv[q];
        k.r[o + i] = si;
        k.x[o + i] = vx[q] + sc.alpha * vX[q];
        smax = nan_max(smax, fabs(si));
        xs[i] = si;
    }
    smax = wave_max(smax);
    if ((threadIdx.x & (kWave - 1)) == 0) wr[threadIdx.x / kWave] = smax;
    __syncthreads();
    if constexpr (PC != CYMG_PRECOND_NONE) apply_block<PC>(k, f, n, xs, k.sh + o, red);
    if (threadIdx.x == 0) {
        double m = wr[0];
        for (int q = 1; q < kApplyWaves; ++q) m = nan_max(m, wr[q]);
        k.part[P_SMAX * k.nfib + f] = m;
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kWave) void k_iter_d(KArgs k)
{
    __shared__ FiberLds<D> s;
        const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const uint32_t sys = k.nsys == 1 ? 0u : a.fib_sys[f];
    const double* X = PC ? k.sh : k.r;
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int n = D * (int)np;
    for (int i = threadIdx.x; i < n; i += kWave) s.x[i] = X[o + i];
    SysScalars sc = k.st_in[sys];
    const bool first = first_of_system(a, sys);
    if (!sc.done) {
        const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
        const double m = wmax(k.part + P_SMAX * k.nfib, fb, fe);
        sc.resid_s = (m != m) ? __builtin_inf() : m;
        sc.stage = ST_D;
    }
    if (first && threadIdx.x == 0) k.st_out[sys] = sc;
    if (sc.done) return;
    load_factors<D>(a, s, a.seg[f], (int)np - 1);
    __syncthreads();
    fiber_operator<D>(a, s, X, f);
    double tt = 0, ts = 0, rts = 0, rtt = 0;
    for (int i = threadIdx.x; i < n; i += kWave) {
        const double ti = s.y[i];
        const double si = k.r[o + i];
        const double rti = k.rt[o + i];
        k.t[o + i] = ti;
        tt += ti * ti;
        ts += ti * si;
        rts += rti * si;
        rtt += rti * ti;
    }
    tt = wave_sum(tt);
    ts = wave_sum(ts);
    rts = wave_sum(rts);
    rtt = wave_sum(rtt);
    if (threadIdx.x == 0)
        *reinterpret_cast<double4*>(k.part4 + 4 * (size_t)f) = make_double4(tt, ts, rts, rtt);
}


// ===========================================================================
// Fused variant: two launches per iteration over tiles of whole fibers.  The
// preconditioned vectors are carried by recurrences of the linear block
// inverse M^-1:
//   y = M^-1 r, z = M^-1 v, w = M^-1 t,
//   phat = M^-1 p = y + beta (phat_old - omega z_old),
//   shat = M^-1 s = y - alpha z,   y_new = M^-1 r_new = shat - omega w,
// so every preconditioned vector a coupling term needs is computed on the fly
// from vectors the current launch does not modify, and the two fiber-local
// phases of the four-launch scheme disappear.  Each launch applies the block
// inverse once, to the operator result it has just computed.
//   P: omega/rho/beta from Q's partials; [x += w shat; r = s - w t; y]; phat;
//      v = M phat; z = M^-1 v; |r|_inf and <rt, v> partials
//   Q: stop test on |r|_inf; alpha; s = r - alpha v; x += alpha phat;
//      shat = y - alpha z; t = M shat; w = M^-1 t; the partials of D
// Partial sums are written per tile and re-reduced, in tile order, by every
// tile of the same system in the next launch.

// storage type and array of the block inverses of a preconditioner mode
template <int PC>
struct PcBlocks {
    using T = double;
    __device__ static const double* get(const KArgs& k) { return PC == CYMG_PRECOND_FP64 ? k.B64 : nullptr; }
};
template <>
struct PcBlocks<CYMG_PRECOND_FP32> {
    using T = float;
    __device__ static const float* get(const KArgs& k) { return k.B32; }
};

// phat(j) for any j, from vectors P does not write
struct PhatOp {
    const double* __restrict__ y;
    const double* __restrict__ sh;
    const double* __restrict__ w;
    const double* __restrict__ ph;
    const double* __restrict__ z;
    double beta, omega;
    int pending;
    __device__ __forceinline__ double yv(uint32_t j) const { return pending ? sh[j] + (-omega) * w[j] : y[j]; }
    __device__ __forceinline__ double operator()(uint32_t j) const
    {
        const double yj = yv(j);
        return beta == 0.0 ? yj : yj + beta * (ph[j] + (-omega) * z[j]);
    }
};

// shat(j) = y(j) - alpha z(j)
struct ShatOp {
    const double* __restrict__ y;
    const double* __restrict__ z;
    double alpha;
    __device__ __forceinline__ double operator()(uint32_t j) const { return y[j] + (-alpha) * z[j]; }
};

__device__ __forceinline__ bool first_tile_of_system(const Tile& T)
{
    return T.t == T.tb;
}

// r = b - M x0 (or b), rt = r, x = x0 (or 0), y = M^-1 r; <rt, r> partials;
// the first tile of each system sets the threshold from the setup partials
template <int D, int PC>
__global__ __launch_bounds__(kTileThreads) void k_tinit(KArgs k, const double* __restrict__ x0, double tolerance,
                                                         int tol_mode, int iter_max, double dt)
{
    constexpr int R = TileRows<D>::R;
    __shared__ TileLds<D> L;
    __shared__ double wr[kTileWaves];
    const FiberArgs& a = k.fa;
    const Tile T = tile_begin<D>(k.ta, blockIdx.x);
    tile_tables<D, PC != CYMG_PRECOND_NONE>(a, L, T, k.blkoff, k.useblk);
    for (int u = threadIdx.x; u < (int)T.nu; u += kTileThreads) L.x[u] = x0 ? x0[T.U0 + u] : 0.0;
    __syncthreads();
    using TB = typename PcBlocks<PC>::T;
    const TB* Bk = PcBlocks<PC>::get(k);
    TileRows<D> rw;
    tile_rows<D>(L, T, rw);
    TilePre<D, TB> pre;
    tile_prefetch<D, TB>(a, L, T, rw, Bk, pre);
    if (x0) tile_operator<D>(a, L, Plain{ x0 }, T, rw, pre);
    double rr = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        if (rw.k[q] < 0) continue;
        const int u = threadIdx.x + q * kTileThreads;
        double ri = k.b[T.U0 + u];
        if (x0) ri = ri + (-1.0) * L.y[u];
        k.r[T.U0 + u] = ri;
        k.rt[T.U0 + u] = ri;
        k.x[T.U0 + u] = L.x[u];
        L.f[u] = ri;
        rr += ri * ri;
    }
    rr = wave_sum(rr);
    if ((threadIdx.x & (kWave - 1)) == 0) wr[threadIdx.x / kWave] = rr;
    __syncthreads();
    tile_block_apply<TB, D>(Bk, L, rw, pre, L.f, k.y + T.U0);
    if (threadIdx.x == 0) {
        double t = wr[0];
        for (int q = 1; q < kTileWaves; ++q) t += wr[q];
        k.part[P_RR * k.ta.ntiles + T.t] = t;
    }
    if (first_tile_of_system(T) && threadIdx.x < kWave) {
        const uint32_t fb = a.sys_begin[T.sys], fe = a.sys_begin[T.sys + 1];
        const double noise = wmin(k.fpart + P_NOISE * k.nfib, fb, fe);
        const double bn = wmax(k.fpart + P_BNORM * k.nfib, fb, fe);
        if (threadIdx.x == 0) {
            SysScalars st{};
            st.noise = noise;
            st.bnorm = bn;
            st.rho1 = 1.0;
            st.alpha = 0.0;
            st.omega = 1.0;
            st.tol = (tol_mode == CYMG_TOL_RELATIVE) ? tolerance * bn : tolerance * (noise * dt);
            st.resid = __builtin_inf();
            st.resid_s = __builtin_inf();
            st.iter = 0;
            st.iter_max = iter_max;
            st.stage = ST_INIT;
            k.st_out[T.sys] = st;
        }
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kTileThreads) void k_tp(KArgs k)
{
    constexpr int R = TileRows<D>::R;
    __shared__ TileLds<D> L;
    __shared__ double wr[2 * kTileWaves];
    __shared__ double bc[5];
    const FiberArgs& a = k.fa;
    const Tile T = tile_begin<D>(k.ta, blockIdx.x);
    double vr[R], vt[R], vs[R], vw[R], vy[R], vph[R], vz[R], vx[R], vrt[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int u = threadIdx.x + q * kTileThreads;
        if (u < (int)T.nu) {
            const uint32_t g = T.U0 + u;
            vr[q] = k.r[g];
            vt[q] = k.t[g];
            vs[q] = k.sh[g];
            vw[q] = k.w[g];
            vy[q] = k.y[g];
            vph[q] = k.ph_in[g];
            vz[q] = k.z_in[g];
            vx[q] = k.x[g];
            vrt[q] = k.rt[g];
        }
    }
    SysScalars sc = k.st_in[T.sys];
    const bool first = first_tile_of_system(T);
    if (sc.done) {
        if (first && threadIdx.x == 0) k.st_out[T.sys] = sc;
        return;
    }
    tile_tables<D, PC != CYMG_PRECOND_NONE>(a, L, T, k.blkoff, k.useblk);
    const uint32_t nt = k.ta.ntiles;
    const uint32_t tb = T.tb, te = T.te;
    sc.pending = 0;
    if (threadIdx.x >= kWave && threadIdx.x < 2 * kWave) {
        if (sc.stage == ST_INIT) {
            const double rr = wsum(k.part + P_RR * nt, tb, te);
            if (threadIdx.x == kWave) bc[0] = rr;
        } else {
            double S4[4];
            wsum4(k.part4, tb, te, S4);
            const double m = wmax(k.part + P_SMAX * nt, tb, te);
            if (threadIdx.x == kWave) { bc[0] = S4[0]; bc[1] = S4[1]; bc[2] = S4[2]; bc[3] = S4[3]; bc[4] = m; }
        }
    }
    __syncthreads();
    if (sc.stage == ST_INIT) {
        sc.rho2 = sc.rho1;
        sc.rho1 = bc[0];
        sc.beta = (sc.rho1 / sc.rho2) * (sc.alpha / sc.omega);
    } else {
        sc.resid_s = (bc[4] != bc[4]) ? __builtin_inf() : bc[4];
        const double tt = bc[0], ts = bc[1], rts = bc[2], rtt = bc[3];
        if (tt == 0.0) {
            sc.done = 1; sc.flag = 0; sc.resid = sc.resid_s;
        } else {
            const double om = ts / tt;
            if (om == 0.0) {
                sc.done = 1; sc.flag = 3; sc.resid = sc.resid_s;
            } else {
                sc.omega = om;
                sc.pending = 1;
                sc.rho2 = sc.rho1;
                sc.rho1 = rts - om * rtt;
                sc.beta = (sc.rho1 / sc.rho2) * (sc.alpha / sc.omega);
                sc.iter += 1;
            }
        }
    }
    sc.stage = ST_A;
    if (first && threadIdx.x == 0) k.st_out[T.sys] = sc;
    if (sc.done) return;
    using TB = typename PcBlocks<PC>::T;
    const TB* Bk = PcBlocks<PC>::get(k);
    TileRows<D> rw;
    tile_rows<D>(L, T, rw);
    TilePre<D, TB> pre;
    tile_prefetch<D, TB>(a, L, T, rw, Bk, pre);
    const PhatOp op{ k.y, k.sh, k.w, k.ph_in, k.z_in, sc.beta, sc.omega, sc.pending };
    double rmax = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        if (rw.k[q] < 0) continue;
        const int u = threadIdx.x + q * kTileThreads;
        const uint32_t g = T.U0 + u;
        double ri = vr[q];
        double yi = vy[q];
        if (sc.pending) {
            k.x[g] = vx[q] + sc.omega * vs[q];
            ri = ri + (-sc.omega) * vt[q];
            k.r[g] = ri;
            yi = vs[q] + (-sc.omega) * vw[q];
            k.y[g] = yi;
        }
        rmax = nan_max(rmax, fabs(ri));
        const double phi = sc.beta == 0.0 ? yi : yi + sc.beta * (vph[q] + (-sc.omega) * vz[q]);
        k.ph_out[g] = phi;
        L.x[u] = phi;
    }
    __syncthreads();
    tile_operator<D>(a, L, op, T, rw, pre);
    double dl = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        if (rw.k[q] < 0) continue;
        const int u = threadIdx.x + q * kTileThreads;
        const uint32_t g = T.U0 + u;
        k.v[g] = L.y[u];
        dl += vrt[q] * L.y[u];
    }
    rmax = wave_max(rmax);
    dl = wave_sum(dl);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        wr[threadIdx.x / kWave] = rmax;
        wr[kTileWaves + threadIdx.x / kWave] = dl;
    }
    tile_block_apply<TB, D>(Bk, L, rw, pre, L.y, k.z_out + T.U0);
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wr[0], d = wr[kTileWaves];
        for (int q = 1; q < kTileWaves; ++q) {
            m = nan_max(m, wr[q]);
            d += wr[kTileWaves + q];
        }
        k.part[P_RMAX * nt + T.t] = m;
        k.part[P_DELTA * nt + T.t] = d;
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kTileThreads) void k_tq(KArgs k)
{
    constexpr int R = TileRows<D>::R;
    __shared__ TileLds<D> L;
    __shared__ double wr[5 * kTileWaves];
    __shared__ double bc[2];
    const FiberArgs& a = k.fa;
    const Tile T = tile_begin<D>(k.ta, blockIdx.x);
    double vr[R], vv[R], vx[R], vph[R], vy[R], vz[R], vrt[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int u = threadIdx.x + q * kTileThreads;
        if (u < (int)T.nu) {
            const uint32_t g = T.U0 + u;
            vr[q] = k.r[g];
            vv[q] = k.v[g];
            vx[q] = k.x[g];
            vph[q] = k.ph_in[g];
            vy[q] = k.y[g];
            vz[q] = k.z_in[g];
            vrt[q] = k.rt[g];
        }
    }
    SysScalars sc = k.st_in[T.sys];
    const bool first = first_tile_of_system(T);
    const uint32_t nt = k.ta.ntiles;
    if (!sc.done) {
        tile_tables<D, PC != CYMG_PRECOND_NONE>(a, L, T, k.blkoff, k.useblk);
        const uint32_t tb = T.tb, te = T.te;
        if (threadIdx.x >= kWave && threadIdx.x < 2 * kWave) {
            const double res = wmax(k.part + P_RMAX * nt, tb, te);
            const double dsum = wsum(k.part + P_DELTA * nt, tb, te);
            if (threadIdx.x == kWave) { bc[0] = res; bc[1] = dsum; }
        }
        __syncthreads();
        decide(sc, bc[0]);
        if (!sc.done) {
            const double delta = bc[1];
            if (PC != CYMG_PRECOND_NONE && delta == 0.0) { sc.done = 1; sc.flag = 4; }
            else sc.alpha = sc.rho1 / delta;
        }
        sc.stage = ST_C;
    }
    if (first && threadIdx.x == 0) k.st_out[T.sys] = sc;
    if (sc.done) return;
    using TB = typename PcBlocks<PC>::T;
    const TB* Bk = PcBlocks<PC>::get(k);
    TileRows<D> rw;
    tile_rows<D>(L, T, rw);
    TilePre<D, TB> pre;
    tile_prefetch<D, TB>(a, L, T, rw, Bk, pre);
    const ShatOp op{ k.y, k.z_in, sc.alpha };
    double smax = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        if (rw.k[q] < 0) continue;
        const int u = threadIdx.x + q * kTileThreads;
        const uint32_t g = T.U0 + u;
        const double si = vr[q] + (-sc.alpha) * vv[q];
        k.r[g] = si;
        k.x[g] = vx[q] + sc.alpha * vph[q];
        smax = nan_max(smax, fabs(si));
        const double shi = vy[q] + (-sc.alpha) * vz[q];
        k.sh[g] = shi;
        L.x[u] = shi;
    }
    __syncthreads();
    tile_operator<D>(a, L, op, T, rw, pre);
    double tt = 0, ts = 0, rts = 0, rtt = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        if (rw.k[q] < 0) continue;
        const int u = threadIdx.x + q * kTileThreads;
        const uint32_t g = T.U0 + u;
        const double ti = L.y[u];
        const double si = vr[q] + (-sc.alpha) * vv[q];
        const double rti = vrt[q];
        k.t[g] = ti;
        tt += ti * ti;
        ts += ti * si;
        rts += rti * si;
        rtt += rti * ti;
    }
    smax = wave_max(smax);
    tt = wave_sum(tt);
    ts = wave_sum(ts);
    rts = wave_sum(rts);
    rtt = wave_sum(rtt);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const int wv = threadIdx.x / kWave;
        wr[wv] = smax;
        wr[kTileWaves + wv] = tt;
        wr[2 * kTileWaves + wv] = ts;
        wr[3 * kTileWaves + wv] = rts;
        wr[4 * kTileWaves + wv] = rtt;
    }
    tile_block_apply<TB, D>(Bk, L, rw, pre, L.y, k.w + T.U0);
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wr[0], s0 = wr[kTileWaves], s1 = wr[2 * kTileWaves], s2 = wr[3 * kTileWaves],
               s3 = wr[4 * kTileWaves];
#pragma unroll
        for (int q = 1; q < kTileWaves; ++q) {
            m = nan_max(m, wr[q]);
            s0 += wr[kTileWaves + q];
            s1 += wr[2 * kTileWaves + q];
            s2 += wr[3 * kTileWaves + q];
            s3 += wr[4 * kTileWaves + q];
        }
        k.part[P_SMAX * nt + T.t] = m;
        *reinterpret_cast<double4*>(k.part4 + 4 * (size_t)T.t) = make_double4(s0, s1, s2, s3);
    }
}

static KArgs make_kargs(Context& c)
{
    KArgs k;
    k.fa = make_fiber_args(c);
    k.blkoff = c.d_blkoff.get();
    k.useblk = c.d_useblk.get();
    k.B64 = c.d_binv.get();
    k.B32 = c.d_binv32.get();
    k.b = c.d_rhs.get();
    k.x = c.d_sol.get();
    k.r = c.k_r.get();
    k.rt = c.k_rt.get();
    k.p = c.k_p.get();
    k.v = c.k_v.get();
    k.t = c.k_t.get();
    k.ph = c.k_ph.get();
    k.sh = c.k_sh.get();
    k.y = c.k_y.get();
    k.w = c.k_w.get();
    k.z_in = k.z_out = k.ph_in = k.ph_out = nullptr;
    k.part = c.d_part.get();
    k.part4 = c.d_part.get() + (size_t)P_TT * c.nfib;
    k.fpart = c.d_part.get();
    k.ta.desc = c.d_tile_desc.get();
    k.ta.ntiles = c.ntiles;
    k.st_in = nullptr;
    k.st_out = nullptr;
    k.nfib = c.nfib;
    k.nsys = c.nsys;
    return k;
}

// State buffers: S0 = d_scal[0:nsys], S1 = d_scal[nsys:2nsys].
// init -> S1; A: S1 -> S0; decide: S0 in place; B: S0 -> S1; C: S1 -> S0; D: S0 -> S1.
template <int D, int PC>
static void run(Context& c, const double* guess, int max_iter, double tol, int tol_mode, cymg_solve_report* rep)
{
    KArgs k = make_kargs(c);
    SysScalars* S0 = c.d_scal.get();
    SysScalars* S1 = S0 + c.nsys;
    auto with = [&](SysScalars* in, SysScalars* out) {
        KArgs q = k;
        q.st_in = in;
        q.st_out = out;
        return q;
    };
    const KArgs kI = with(S0, S1), kA = with(S1, S0), kB = with(S0, S1), kC = with(S1, S0), kD = with(S0, S1);
    const KArgs kDec = with(S0, S0);
    const dim3 g(c.nfib), b(kWave), b4(kApplyThreads), gs(c.nsys);
    hipLaunchKernelGGL((k_init<D, PC>), g, b, 0, c.stream, kI, guess, tol, tol_mode, max_iter, c.prm.time_step);
    hipLaunchKernelGGL((k_iter_a<D, PC>), g, b4, 0, c.stream, kA);
    hipLaunchKernelGGL(k_decide, gs, b, 0, c.stream, kDec, c.nsys);
    CYMG_HIP(hipGetLastError());
    // the first chunk covers the iteration count of the previous solve in this
    // mode, so a typical solve needs one or two host reads of the stop flag;
    // iterations launched after convergence exit at once
    int launched = 0;
    int& predicted = c.last_iters[PC];
    int chunk = predicted > 0 ? predicted + 1 : 4;
    const size_t sbytes = sizeof(SysScalars) * c.nsys;
    while (true) {
        CYMG_HIP(hipMemcpyAsync(c.h_scal, S0, sbytes, hipMemcpyDeviceToHost, c.stream));
        CYMG_HIP(hipStreamSynchronize(c.stream));
        bool all = true;
        for (uint32_t s = 0; s < c.nsys; ++s) all = all && c.h_scal[s].done;
        if (all || launched > max_iter + 2) break;
        for (int it = 0; it < chunk; ++it) {
            hipLaunchKernelGGL((k_iter_b<D, PC>), g, b, 0, c.stream, kB);
            hipLaunchKernelGGL((k_iter_c<D, PC>), g, b4, 0, c.stream, kC);
            hipLaunchKernelGGL((k_iter_d<D, PC>), g, b, 0, c.stream, kD);
            hipLaunchKernelGGL((k_iter_a<D, PC>), g, b4, 0, c.stream, kA);
        }
        hipLaunchKernelGGL(k_decide, gs, b, 0, c.stream, kDec, c.nsys);
        CYMG_HIP(hipGetLastError());
        launched += chunk;
        chunk = predicted > 0 ? std::max(4, predicted / 8) : std::min(2 * chunk, 16);
    }
    // per-system reports; the summary is the worst system (most iterations,
    // any failure), system 0 when there is one system
    c.sys_reports.assign(c.nsys, cymg_solve_report{});
    int worst = 0, all_conv = 1;
    for (uint32_t s = 0; s < c.nsys; ++s) {
        const SysScalars& w = c.h_scal[s];
        cymg_solve_report& r = c.sys_reports[s];
        r.iterations = w.iter;
        r.total_iterations = w.iter;
        r.flag = w.flag;
        r.residual = w.resid;
        r.threshold = w.tol;
        r.converged = w.resid < w.tol;
        r.noise_level = w.noise * c.prm.time_step;
        r.precond_used = PC;
        all_conv = all_conv && r.converged;
        if (w.iter > c.h_scal[worst].iter || (!r.converged && c.sys_reports[worst].converged)) worst = (int)s;
    }
    *rep = c.sys_reports[worst];
    rep->converged = all_conv;
    predicted = c.h_scal[worst].iter;
}

template <int D, int PC>
static void run_tiled(Context& c, const double* guess, int max_iter, double tol, int tol_mode,
                      cymg_solve_report* rep)
{
    KArgs k = make_kargs(c);
    k.part = c.d_tpart.get();
    k.part4 = c.d_tpart.get() + (size_t)P_TT * c.ntiles;
    SysScalars* S0 = c.d_scal.get();
    SysScalars* S1 = S0 + c.nsys;
    double* ph[2] = { c.k_ph.get(), c.k_ph2.get() };
    double* zz[2] = { c.k_z.get(), c.k_z2.get() };
    auto with = [&](SysScalars* in, SysScalars* out, int par) {
        KArgs q = k;
        q.st_in = in;
        q.st_out = out;
        q.ph_in = ph[par];
        q.ph_out = ph[par ^ 1];
        q.z_in = zz[par];
        q.z_out = zz[par ^ 1];
        return q;
    };
    // P reads phat/z of its parity and writes the other; Q reads what P wrote
    const KArgs kI = with(S0, S1, 0);
    const KArgs kP0 = with(S1, S0, 0), kQ0 = with(S0, S1, 1);
    const KArgs kP1 = with(S1, S0, 1), kQ1 = with(S0, S1, 0);
    const dim3 g(c.ntiles), b(kTileThreads);
    hipLaunchKernelGGL((k_tinit<D, PC>), g, b, 0, c.stream, kI, guess, tol, tol_mode, max_iter, c.prm.time_step);
    CYMG_HIP(hipGetLastError());
    int launched = 0;
    int& predicted = c.last_iters[PC];
    int chunk = predicted > 0 ? predicted + 2 : 4;
    const size_t sbytes = sizeof(SysScalars) * c.nsys;
    bool first = true;
    while (true) {
        if (!first) {
            CYMG_HIP(hipMemcpyAsync(c.h_scal, S1, sbytes, hipMemcpyDeviceToHost, c.stream));
            CYMG_HIP(hipStreamSynchronize(c.stream));
            bool all = true;
            for (uint32_t s = 0; s < c.nsys; ++s) all = all && c.h_scal[s].done;
            if (all || launched > max_iter + 4) break;
        }
        first = false;
        for (int it = 0; it < chunk; it += 2) {
            hipLaunchKernelGGL((k_tp<D, PC>), g, b, 0, c.stream, kP0);
            hipLaunchKernelGGL((k_tq<D, PC>), g, b, 0, c.stream, kQ0);
            hipLaunchKernelGGL((k_tp<D, PC>), g, b, 0, c.stream, kP1);
            hipLaunchKernelGGL((k_tq<D, PC>), g, b, 0, c.stream, kQ1);
        }
        CYMG_HIP(hipGetLastError());
        launched += chunk;
        chunk = predicted > 0 ? std::max(4, predicted / 8) : std::min(2 * chunk, 16);
        chunk += chunk & 1;
    }
    c.sys_reports.assign(c.nsys, cymg_solve_report{});
    int worst = 0, all_conv = 1;
    for (uint32_t s = 0; s < c.nsys; ++s) {
        const SysScalars& w = c.h_scal[s];
        cymg_solve_report& r = c.sys_reports[s];
        r.iterations = w.iter;
        r.total_iterations = w.iter;
        r.flag = w.flag;
        r.residual = w.resid;
        r.threshold = w.tol;
        r.converged = w.resid < w.tol;
        r.noise_level = w.noise * c.prm.time_step;
        r.precond_used = PC;
        all_conv = all_conv && r.converged;
        if (w.iter > c.h_scal[worst].iter || (!r.converged && c.sys_reports[worst].converged)) worst = (int)s;
    }
    *rep = c.sys_reports[worst];
    rep->converged = all_conv;
    predicted = c.h_scal[worst].iter;
}

// Average duration of one launch of the two-launch iteration (P and Q
// alternately), from events around `pairs` iterations on the context's
// stream.  The iteration runs with a zero threshold so that every launch does
// its full work; an iteration that stops early (breakdown) is an error.
template <int D, int PC>
static double time_tiled(Context& c, int pairs)
{
    KArgs k = make_kargs(c);
    k.part = c.d_tpart.get();
    k.part4 = c.d_tpart.get() + (size_t)P_TT * c.ntiles;
    SysScalars* S0 = c.d_scal.get();
    SysScalars* S1 = S0 + c.nsys;
    double* ph[2] = { c.k_ph.get(), c.k_ph2.get() };
    double* zz[2] = { c.k_z.get(), c.k_z2.get() };
    auto with = [&](SysScalars* in, SysScalars* out, int par) {
        KArgs q = k;
        q.st_in = in;
        q.st_out = out;
        q.ph_in = ph[par];
        q.ph_out = ph[par ^ 1];
        q.z_in = zz[par];
        q.z_out = zz[par ^ 1];
        return q;
    };
    const KArgs kI = with(S0, S1, 0);
    const KArgs kP0 = with(S1, S0, 0), kQ0 = with(S0, S1, 1);
    const KArgs kP1 = with(S1, S0, 1), kQ1 = with(S0, S1, 0);
    const dim3 g(c.ntiles), b(kTileThreads);
    hipLaunchKernelGGL((k_tinit<D, PC>), g, b, 0, c.stream, kI, nullptr, 0.0, CYMG_TOL_RELATIVE, 1 << 30,
                       c.prm.time_step);
    pairs += pairs & 1;
    CYMG_HIP(hipEventRecord(c.ev[8], c.stream));
    for (int it = 0; it < pairs; it += 2) {
        hipLaunchKernelGGL((k_tp<D, PC>), g, b, 0, c.stream, kP0);
        hipLaunchKernelGGL((k_tq<D, PC>), g, b, 0, c.stream, kQ0);
        hipLaunchKernelGGL((k_tp<D, PC>), g, b, 0, c.stream, kP1);
        hipLaunchKernelGGL((k_tq<D, PC>), g, b, 0, c.stream, kQ1);
    }
    CYMG_HIP(hipGetLastError());
    CYMG_HIP(hipEventRecord(c.ev[9], c.stream));
    CYMG_HIP(hipEventSynchronize(c.ev[9]));
    float ms = 0;
    CYMG_HIP(hipEventElapsedTime(&ms, c.ev[8], c.ev[9]));
    CYMG_HIP(hipMemcpyAsync(c.h_scal, S1, sizeof(SysScalars) * c.nsys, hipMemcpyDeviceToHost, c.stream));
    CYMG_HIP(hipStreamSynchronize(c.stream));
    for (uint32_t s = 0; s < c.nsys; ++s)
        if (c.h_scal[s].done) throw Error(CYMG_ERR_INVALID, "the timed iteration stopped early; use fewer launches");
    return ms / (2.0 * pairs);
}

static void tiled_buffers(Context& c)
{
    const size_t n = c.n();
    const bool fresh = c.k_y.cap < n;
    c.d_tpart.resize((size_t)P_SLOTS * (c.ntiles ? c.ntiles : 1));
    c.k_r.resize(n); c.k_rt.resize(n); c.k_v.resize(n); c.k_t.resize(n); c.k_sh.resize(n);
    c.k_ph.resize(n); c.k_ph2.resize(n); c.k_y.resize(n); c.k_z.resize(n); c.k_z2.resize(n); c.k_w.resize(n);
    if (fresh)
        for (DevBuf<double>* v : { &c.k_ph, &c.k_ph2, &c.k_z, &c.k_z2, &c.k_w, &c.k_sh, &c.k_t })
            CYMG_HIP(hipMemsetAsync(v->get(), 0, v->cap * sizeof(double), c.stream));
}

double krylov_time_launch(Context& c, int precond, int launches)
{
    tiled_buffers(c);
    const int pairs = std::max(1, launches / 2);
    if (c.dim == 3) {
        if (precond == CYMG_PRECOND_FP64) return time_tiled<3, CYMG_PRECOND_FP64>(c, pairs);
        if (precond == CYMG_PRECOND_FP32) return time_tiled<3, CYMG_PRECOND_FP32>(c, pairs);
        return time_tiled<3, CYMG_PRECOND_NONE>(c, pairs);
    }
    if (precond == CYMG_PRECOND_FP64) return time_tiled<2, CYMG_PRECOND_FP64>(c, pairs);
    if (precond == CYMG_PRECOND_FP32) return time_tiled<2, CYMG_PRECOND_FP32>(c, pairs);
    return time_tiled<2, CYMG_PRECOND_NONE>(c, pairs);
}

void krylov_solve(Context& c, int precond, const double* guess_dev, int max_iter, double tol, int tol_mode,
                  cymg_solve_report* rep)
{
    if (c.krylov_fused) {
        tiled_buffers(c);
        if (c.dim == 3) {
            if (precond == CYMG_PRECOND_FP64) run_tiled<3, CYMG_PRECOND_FP64>(c, guess_dev, max_iter, tol, tol_mode, rep);
            else if (precond == CYMG_PRECOND_FP32) run_tiled<3, CYMG_PRECOND_FP32>(c, guess_dev, max_iter, tol, tol_mode, rep);
            else run_tiled<3, CYMG_PRECOND_NONE>(c, guess_dev, max_iter, tol, tol_mode, rep);
        } else {
            if (precond == CYMG_PRECOND_FP64) run_tiled<2, CYMG_PRECOND_FP64>(c, guess_dev, max_iter, tol, tol_mode, rep);
            else if (precond == CYMG_PRECOND_FP32) run_tiled<2, CYMG_PRECOND_FP32>(c, guess_dev, max_iter, tol, tol_mode, rep);
            else run_tiled<2, CYMG_PRECOND_NONE>(c, guess_dev, max_iter, tol, tol_mode, rep);
        }
        return;
    }
    const size_t n = c.n();
    c.k_r.resize(n); c.k_rt.resize(n); c.k_p.resize(n); c.k_v.resize(n); c.k_t.resize(n);
    c.k_ph.resize(n); c.k_sh.resize(n);
    if (c.dim == 3) {
        if (precond == CYMG_PRECOND_FP64) run<3, CYMG_PRECOND_FP64>(c, guess_dev, max_iter, tol, tol_mode, rep);
        else if (precond == CYMG_PRECOND_FP32) run<3, CYMG_PRECOND_FP32>(c, guess_dev, max_iter, tol, tol_mode, rep);
        else run<3, CYMG_PRECOND_NONE>(c, guess_dev, max_iter, tol, tol_mode, rep);
    } else {
        if (precond == CYMG_PRECOND_FP64) run<2, CYMG_PRECOND_FP64>(c, guess_dev, max_iter, tol, tol_mode, rep);
        else if (precond == CYMG_PRECOND_FP32) run<2, CYMG_PRECOND_FP32>(c, guess_dev, max_iter, tol, tol_mode, rep);
        else run<2, CYMG_PRECOND_NONE>(c, guess_dev, max_iter, tol, tol_mode, rep);
    }
}

}  // namespace cymg
