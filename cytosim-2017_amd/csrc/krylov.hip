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
#include <cstring>
#include <type_traits>

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
    // two-launch iteration: records gathered at coupled points
    double4* g1a;       // (shat, w, phat, z) of the previous iteration: read by P
    double4* g1b;       // the same for this iteration: P writes (phat, z), Q writes (shat, w)
    double2* g2;        // (y, z) of this iteration: P writes, Q reads
    double* part;       // partial sums [slot][fiber] (four launches) or [slot][tile] (two)
    double* part4;      // [fiber or tile][4] partials of D
    const double* fpart;// per-fiber partials of the setup (noise, |b|_inf)
    TileArgs ta;
    SysScalars* st_in;
    SysScalars* st_out;
    uint32_t nfib;      // Mecables: the stride of the partial sums (fibers, then Solids)
    uint32_t nsys;
    uint32_t nfiber;    // fibers: workgroups from nfiber on handle Solid blockIdx.x - nfiber
    int f32apply;       // FP32 blocks applied in FP32 arithmetic (four-launch iteration only)
    SolidArgs sa;
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
    return wave_min(v);
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

// The scalar state is read and written field by field: a whole-struct copy
// lets the compiler keep part of it in a per-thread LDS array indexed by the
// flat work-item id, which makes every wavefront read the workgroup size from
// the dispatch packet in host memory at the start of the kernel.
__device__ __forceinline__ SysScalars state_load(const SysScalars* __restrict__ p)
{
    SysScalars s;
    s.rho1 = p->rho1; s.rho2 = p->rho2; s.alpha = p->alpha; s.omega = p->omega; s.beta = p->beta;
    s.tol = p->tol; s.resid = p->resid; s.resid_s = p->resid_s; s.bnorm = p->bnorm; s.noise = p->noise;
    s.iter = p->iter; s.iter_max = p->iter_max; s.done = p->done; s.flag = p->flag;
    s.pending = p->pending; s.stage = p->stage;
    return s;
}

__device__ __forceinline__ void state_store(SysScalars* __restrict__ p, const SysScalars& s)
{
    p->rho1 = s.rho1; p->rho2 = s.rho2; p->alpha = s.alpha; p->omega = s.omega; p->beta = s.beta;
    p->tol = s.tol; p->resid = s.resid; p->resid_s = s.resid_s; p->bnorm = s.bnorm; p->noise = s.noise;
    p->iter = s.iter; p->iter_max = s.iter_max; p->done = s.done; p->flag = s.flag;
    p->pending = s.pending; p->stage = s.stage;
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
    else if constexpr (PC == CYMG_PRECOND_FP32) {
        if (k.f32apply) block_apply_mw_f32<kApplyWaves>(k.B32 + k.blkoff[f], bs, k.useblk[f] != 0, xs, y, red);
        else block_apply_mw<float, kApplyWaves>(k.B32 + k.blkoff[f], bs, k.useblk[f] != 0, xs, y, red);
    }
}

// row extents a Solid workgroup of the operator phases stages in the
// FiberLds force array it does not otherwise use
template <int D>
constexpr uint32_t kLdsRows = 2 * kMaxFiberPoints * D;

// ---------------------------------------------------------------------------
// The rows of Solid sk in each phase (the same vector updates as the fibers'
// rows, with the Solid operator and its low-rank block inverse); red: LDS
// scratch of at least 96 doubles.

template <int D>
__device__ double solid_init(const KArgs& k, uint32_t sk, const double* __restrict__ x0, double* red, uint32_t* lrow,
                             uint32_t cap)
{
    const SolidArgs& s = k.sa;
    const uint32_t o = D * s.idx[sk], n = D * s.np[sk];
    if (x0) {
        double* r = k.r;
        const double* b = k.b;
        solid_operator<D>(k.fa, s, sk, Plain{ x0 }, [=](uint32_t i, double y) { r[i] = b[i] + (-1.0) * y; }, red,
                          lrow, cap);
        __syncthreads();
    }
    double rr = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double ri = x0 ? k.r[o + i] : k.b[o + i];
        k.r[o + i] = ri;
        k.rt[o + i] = ri;
        k.x[o + i] = x0 ? x0[o + i] : 0.0;
        rr += ri * ri;
    }
    double v[1] = { rr };
    wg_sum<1>(v, red);
    return v[0];
}

// A: [x += omega shat; r = s - omega t]; p = r + beta (p - omega v); phat = M^-1 p
template <int D, int PC>
__device__ void solid_iter_a(const KArgs& k, uint32_t sk, const SysScalars& sc, double* red)
{
    const SolidArgs& s = k.sa;
    const uint32_t o = D * s.idx[sk], n = D * s.np[sk];
    double rmax = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        double ri = k.r[o + i];
        if (sc.pending) {
            const double si = PC ? k.sh[o + i] : ri;
            k.x[o + i] = k.x[o + i] + sc.omega * si;
            ri = ri + (-sc.omega) * k.t[o + i];
            k.r[o + i] = ri;
        }
        rmax = nan_max(rmax, fabs(ri));
        double pi;
        if (sc.beta == 0.0) pi = ri;
        else pi = sc.beta * (k.p[o + i] + (-sc.omega) * k.v[o + i]) + ri;
        k.p[o + i] = pi;
    }
    rmax = wg_max(rmax, red);
    if constexpr (PC != CYMG_PRECOND_NONE) {
        __syncthreads();
        double* ph = k.ph;
        solid_precond<D>(s, sk, Plain{ k.p }, [=](uint32_t i, double y) { ph[i] = y; }, red);
    }
    if (threadIdx.x == 0) k.part[P_RMAX * k.nfib + k.nfiber + sk] = rmax;
}

// B: v = M phat; <rt, v>
template <int D, int PC>
__device__ void solid_iter_b(const KArgs& k, uint32_t sk, double* red, uint32_t* lrow, uint32_t cap)
{
    const double* X = PC ? k.ph : k.p;
    double* v = k.v;
    const double* rt = k.rt;
    double dl = 0.0;
    solid_operator<D>(k.fa, k.sa, sk, Plain{ X }, [&](uint32_t i, double y) {
        v[i] = y;
        dl += rt[i] * y;
    }, red, lrow, cap);
    double w[1] = { dl };
    wg_sum<1>(w, red);
    if (threadIdx.x == 0) k.part[P_DELTA * k.nfib + k.nfiber + sk] = w[0];
}

// C: s = r - alpha v; x += alpha phat; shat = M^-1 s
template <int D, int PC>
__device__ void solid_iter_c(const KArgs& k, uint32_t sk, const SysScalars& sc, double* red)
{
    const SolidArgs& s = k.sa;
    const uint32_t o = D * s.idx[sk], n = D * s.np[sk];
    const double* X = PC ? k.ph : k.p;
    double smax = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double si = k.r[o + i] + (-sc.alpha) * k.v[o + i];
        k.r[o + i] = si;
        k.x[o + i] = k.x[o + i] + sc.alpha * X[o + i];
        smax = nan_max(smax, fabs(si));
    }
    smax = wg_max(smax, red);
    if constexpr (PC != CYMG_PRECOND_NONE) {
        __syncthreads();
        double* sh = k.sh;
        solid_precond<D>(s, sk, Plain{ k.r }, [=](uint32_t i, double y) { sh[i] = y; }, red);
    }
    if (threadIdx.x == 0) k.part[P_SMAX * k.nfib + k.nfiber + sk] = smax;
}

// D: t = M shat; <t,t>, <t,s>, <rt,s>, <rt,t>
template <int D, int PC>
__device__ void solid_iter_d(const KArgs& k, uint32_t sk, double* red, uint32_t* lrow, uint32_t cap)
{
    const double* X = PC ? k.sh : k.r;
    double* t = k.t;
    const double* sv = k.r;
    const double* rt = k.rt;
    double acc[4] = { 0, 0, 0, 0 };
    solid_operator<D>(k.fa, k.sa, sk, Plain{ X }, [&](uint32_t i, double y) {
        t[i] = y;
        acc[0] += y * y;
        acc[1] += y * sv[i];
        acc[2] += rt[i] * sv[i];
        acc[3] += rt[i] * y;
    }, red, lrow, cap);
    wg_sum<4>(acc, red);
    if (threadIdx.x == 0)
        *reinterpret_cast<double4*>(k.part4 + 4 * (size_t)(k.nfiber + sk)) = make_double4(acc[0], acc[1], acc[2], acc[3]);
}

// r = b - M x0 (or b), rt = r, x = x0 (or 0); <rt, r> partials; threshold
template <int D, int PC>
__global__ __launch_bounds__(kWave) void k_init(KArgs k, const double* __restrict__ x0, double tolerance,
                                                int tol_mode, int iter_max, double dt)
{
    __shared__ FiberLds<D> s;
    const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    double rr = 0.0;
    if (f >= k.nfiber) {
        rr = solid_init<D>(k, f - k.nfiber, x0, s.x, reinterpret_cast<uint32_t*>(s.f), kLdsRows<D>);
    } else {
        const uint32_t np = a.np[f], o = D * a.idx[f];
        const int n = D * (int)np;
        if (x0) {
            for (int i = threadIdx.x; i < n; i += kWave) s.x[i] = x0[o + i];
            load_factors<D>(a, s, a.seg[f], (int)np - 1);
            __syncthreads();
            fiber_operator<D>(a, s, x0, f);
        }
        for (int i = threadIdx.x; i < n; i += kWave) {
            double ri = k.b[o + i];
            if (x0) ri = ri + (-1.0) * s.y[i];
            k.r[o + i] = ri;
            k.rt[o + i] = ri;
            k.x[o + i] = x0 ? s.x[i] : 0.0;
            rr += ri * ri;
        }
        rr = wave_sum(rr);
    }
    if (threadIdx.x == 0) k.part[P_RR * k.nfib + f] = rr;
    const uint32_t sys = (k.nsys == 1 || f >= k.nfiber) ? 0u : a.fib_sys[f];
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
            state_store(k.st_out + sys, w);
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
    const bool solid = f >= k.nfiber;
    const uint32_t sys = (k.nsys == 1 || solid) ? 0u : a.fib_sys[f];
    const uint32_t np = solid ? 0u : a.np[f], o = solid ? 0u : D * a.idx[f];
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
    SysScalars sc = state_load(k.st_in + sys);
    const bool first = first_of_system(a, sys);
    if (sc.done) {
        if (first && threadIdx.x == 0) state_store(k.st_out + sys, sc);
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
    if (first && threadIdx.x == 0) state_store(k.st_out + sys, sc);
    if (sc.done) return;
    if (solid) {
        solid_iter_a<D, PC>(k, f - k.nfiber, sc, red);
        return;
    }
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
    SysScalars sc = state_load(k.st_in + sys);
    if (sc.done || sc.stage != ST_A) return;
    const uint32_t fb = k.fa.sys_begin[sys], fe = k.fa.sys_begin[sys + 1];
    const double res = wmax(k.part + P_RMAX * k.nfib, fb, fe);
    if (threadIdx.x == 0) {
        decide(sc, res);
        state_store(k.st_in + sys, sc);
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kWave) void k_iter_b(KArgs k)
{
    __shared__ FiberLds<D> s;
        const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const bool solid = f >= k.nfiber;
    const uint32_t sys = (k.nsys == 1 || solid) ? 0u : a.fib_sys[f];
    // the operand slice is loaded while the scalar state is derived
    const double* X = PC ? k.ph : k.p;
    const uint32_t np = solid ? 0u : a.np[f], o = solid ? 0u : D * a.idx[f];
    const int n = D * (int)np;
    for (int i = threadIdx.x; i < n; i += kWave) s.x[i] = X[o + i];
    SysScalars sc = state_load(k.st_in + sys);
    const bool first = first_of_system(a, sys);
    if (!sc.done) {
        const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
        decide(sc, wmax(k.part + P_RMAX * k.nfib, fb, fe));
        sc.stage = ST_B;
    }
    if (first && threadIdx.x == 0) state_store(k.st_out + sys, sc);
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
    const bool solid = f >= k.nfiber;
    const uint32_t sys = (k.nsys == 1 || solid) ? 0u : a.fib_sys[f];
    const double* X = PC ? k.ph : k.p;
    const uint32_t np = solid ? 0u : a.np[f], o = solid ? 0u : D * a.idx[f];
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
    SysScalars sc = state_load(k.st_in + sys);
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
    if (first && threadIdx.x == 0) state_store(k.st_out + sys, sc);
    if (sc.done) return;
    if (solid) {
        solid_iter_c<D, PC>(k, f - k.nfiber, sc, red);
        return;
    }
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
    const bool solid = f >= k.nfiber;
    const uint32_t sys = (k.nsys == 1 || solid) ? 0u : a.fib_sys[f];
    const double* X = PC ? k.sh : k.r;
    const uint32_t np = solid ? 0u : a.np[f], o = solid ? 0u : D * a.idx[f];
    const int n = D * (int)np;
    for (int i = threadIdx.x; i < n; i += kWave) s.x[i] = X[o + i];
    SysScalars sc = state_load(k.st_in + sys);
    const bool first = first_of_system(a, sys);
    if (!sc.done) {
        const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
        const double m = wmax(k.part + P_SMAX * k.nfib, fb, fe);
        sc.resid_s = (m != m) ? __builtin_inf() : m;
        sc.stage = ST_D;
    }
    if (first && threadIdx.x == 0) state_store(k.st_out + sys, sc);
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


// The operator phases B and D of the Solids, launched after the fibers' phase
// (same scalar state in, partial sums in the Solids' slots): a Solid's rows
// hold many more coupling entries than a fiber's (one link per aster fiber on
// the central point), so each Solid gets a workgroup of kSolidOpThreads.  The
// fibers' launch writes the phase's state; a Solid workgroup writes it only
// when there is no fiber.
constexpr int kSolidOpThreads = 1024;
constexpr uint32_t kSolidOpRows = 4096;

template <int D, int PC, int PHASE>
__global__ __launch_bounds__(kSolidOpThreads) void k_solid_phase(KArgs k)
{
    __shared__ double red[16 * 6];
    __shared__ uint32_t lrow[kSolidOpRows];
    const uint32_t sk = blockIdx.x;
    SysScalars sc = state_load(k.st_in);
    const bool first = k.nfiber == 0 && sk == 0;
    if (!sc.done) {
        if constexpr (PHASE == ST_B) {
            decide(sc, wmax(k.part + P_RMAX * k.nfib, 0, k.nfib));
        } else {
            const double m = wmax(k.part + P_SMAX * k.nfib, 0, k.nfib);
            sc.resid_s = (m != m) ? __builtin_inf() : m;
        }
        sc.stage = PHASE;
    }
    if (first && threadIdx.x == 0) state_store(k.st_out, sc);
    if (sc.done) return;
    if constexpr (PHASE == ST_B) solid_iter_b<D, PC>(k, sk, red, lrow, kSolidOpRows);
    else solid_iter_d<D, PC>(k, sk, red, lrow, kSolidOpRows);
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

// ---------------------------------------------------------------------------
// Solid tiles of the two-launch iteration: a tile without fibers (nft == 0)
// holds one Solid (Mecable fb), after the fiber tiles of its system.  Its
// rows follow the recurrences of the fiber rows; the operator's Solid rows
// are the operand plus the rigid motion that the moments of the coupling
// forces produce (solid_operator) and the block inverse is I - G W
// (solid_precond).  Thread t owns the unknowns t, t + blockDim.x, ...; the
// operand at coupled unknowns outside the Solid comes from their records, as
// in the fiber tiles.  The per-launch partial sums go to the tile's slots.
constexpr uint32_t kSolidTileRows = 1024;

struct SolidTileLds {
    double a[kSolidTileRows];       // operand of the Solid's rows
    double y[kSolidTileRows];       // operator result
    double e[kSolidTileRows];       // a row value needed with the block apply's output
    double red[16 * 6];
    double bc[8];
    double wr[5 * 16];
    uint32_t lrow[kSolidTileRows + 1];
};

// k_tinit for a Solid tile
template <int D, int PC>
__device__ void solid_tile_init(const KArgs& k, const Tile& T, const double* __restrict__ x0, SolidTileLds& S)
{
    const uint32_t ks = T.fb - k.nfiber, U0 = T.U0, nu = T.nu;
    const auto X0 = [&](uint32_t i) -> double { return x0 ? x0[i] : 0.0; };
    if (x0) {
        const auto O = [&](uint32_t i, double v) { S.y[i - U0] = v; };
        solid_operator<D>(k.fa, k.sa, ks, X0, O, S.red, S.lrow, kSolidTileRows + 1);
        __syncthreads();
    }
    double rr = 0.0;
    for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x) {
        const uint32_t g = U0 + u;
        double ri = k.b[g];
        if (x0) ri = ri + (-1.0) * S.y[u];
        k.r[g] = ri;
        k.rt[g] = ri;
        k.x[g] = X0(g);
        S.e[u] = ri;
        rr += ri * ri;
    }
    rr = wave_sum(rr);
    if ((threadIdx.x & (kWave - 1)) == 0) S.wr[threadIdx.x / kWave] = rr;
    __syncthreads();
    double* y0 = reinterpret_cast<double*>(k.g2);     // (y, z) records: y = M^-1 r
    if (PC == CYMG_PRECOND_NONE) {
        for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x) y0[2 * (size_t)(U0 + u)] = S.e[u];
    } else {
        const auto In = [&](uint32_t i) -> double { return S.e[i - U0]; };
        const auto Out = [&](uint32_t i, double v) { y0[2 * (size_t)i] = v; };
        solid_precond<D>(k.sa, ks, In, Out, S.red);
    }
    if (threadIdx.x == 0) {
        double t = S.wr[0];
        for (uint32_t q = 1; q < blockDim.x / kWave; ++q) t += S.wr[q];
        k.part[P_RR * k.ta.ntiles + T.t] = t;
    }
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
    __shared__ SolidTileLds SL;
    const FiberArgs& a = k.fa;
    const Tile T = tile_begin<D>(k.ta, blockIdx.x);
    if (T.nft == 0) {
        solid_tile_init<D, PC>(k, T, x0, SL);
        return;
    }
    tile_tables<PC != CYMG_PRECOND_NONE>(a, L, T, k.blkoff, k.useblk);
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
    tile_block_apply<TB, D>(Bk, L, rw, pre, L.f, reinterpret_cast<double*>(k.g2 + T.U0), 2);
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
            state_store(k.st_out + T.sys, st);
        }
    }
}

// ---------------------------------------------------------------------------
// P and Q run NT threads per tile: 512, or 256 on layouts without Solid
// tiles, whose smaller workgroup and LDS let two tiles share a compute unit so
// that one tile's loads overlap the other's dependent phases (the long,
// HBM-bound launches of many tiles).  A tile of at most NT/2 unknowns
// gives each row two threads (h = 0 owns the row, h = 1 helps): each takes
// half of the row's mB and mC entries and half of the columns of its row of
// the block inverse.  Everything that does not depend on the iteration
// scalars is requested in the prologue: the tile descriptor, then the row's
// vectors, CSR extents and entries, the raw vectors at the coupled points
// (the operand of another tile is a combination of vectors the launch does
// not write), P' and rfDiff constants; after the first barrier the block
// inverse columns and the segment constants.  A longer tile (one fiber of
// more than NT/2 / DIM points) gives each row one thread.
constexpr int kTpThreads = 512;        // the workgroup of layouts with Solid tiles
constexpr int kTpPreB = 6;      // mB entries per thread in registers
constexpr int kTpPreC = 3;      // mC entries per thread in registers
// 256-thread tiles prefetch less per thread (8 block columns, 2 mB and 1 mC
// entries): at most 128 VGPRs, so four tiles share a compute unit and one
// tile's loads overlap the others' dependent phases.  On the 64-system
// workload this takes a full-work launch from 647 to 522 us (FP64) and from
// 595 to 394 us (FP32); 16/3/2 and 16/2/1 (3 tiles per compute unit), or
// batches of 16 or 8 columns, measured slower.
constexpr int kTp256Blk = 8, kTp256PreB = 2, kTp256PreC = 1;
// Staged launches: the 512 threads of the tile plus kTpStageWaves wavefronts
// that copy the tile's block inverses into LDS (global-to-LDS loads, no
// registers) from the start of the launch, while the other wavefronts run the
// prologue, the gathers and the operator; the block apply then reads the
// staged columns from LDS.  Columns beyond the stage come from memory as in
// the other launches, the first kTpStagedBlk of each thread's share
// prefetched into registers.  Twelve wavefronts leave 168 VGPRs per thread
// (three per SIMD): fewer coupling entries are prefetched than with 512
// threads.  On the 1000-fiber system the FP32 blocks' launch takes 13.4
// instead of 14.0-14.2 us; the FP64 blocks' 16.4 instead of 15.1 us (their
// 112 KB per tile arrive too late), so only FP32 blocks are staged by default.
constexpr int kTpStageWaves = 4;
constexpr int kTpStaged = kTpThreads + kTpStageWaves * kWave;
constexpr int kTpStagedBlk = 8, kTpStagedPreB = 5, kTpStagedPreC = 2;
constexpr uint32_t kStageBytes = 128 * 1024;

template <int NT>
struct TpCfg {
    static constexpr int work = NT == kTpStaged ? kTpThreads : NT;   // threads that own unknowns
    static constexpr int waves = work / kWave;
    static constexpr int split = work / 2;                // unknowns a split tile may hold
    // block-inverse columns per thread in registers
    static constexpr int blk = NT == kTpStaged ? kTpStagedBlk : NT == 512 ? 48 : kTp256Blk;
    static constexpr int preb = NT == kTpStaged ? kTpStagedPreB : NT == 512 ? kTpPreB : kTp256PreB;
    static constexpr int prec = NT == kTpStaged ? kTpStagedPreC : NT == 512 ? kTpPreC : kTp256PreC;
    static constexpr int chunk = kApplyChunk;                 // block columns per load batch beyond blk
    static __host__ __device__ constexpr int max_points(int D)
    {
        return work / D < kTileMaxPoints ? work / D : kTileMaxPoints;
    }
};

template <int D, int NT>
struct TpLds {
    static constexpr int MP = TpCfg<NT>::max_points(D);
    double x[MP * D];      // operand slice
    double f[MP * D];      // forces
    double y[MP * D];      // operator result
    double hb[MP * D];     // helper's mB partial sum, then its block partial sum
    double hc[MP * D];     // helper's mC partial sum
    double lam[MP];
    double pa[MP], pb[MP], pc[MP];
    double fk[kTileMaxFibers];
    double fsca[kTileMaxFibers];
    uint64_t fblk[kTileMaxFibers];
    uint32_t fo[kTileMaxFibers + 1];
    uint32_t fseg[kTileMaxFibers];
    int fuse[kTileMaxFibers];
    uint32_t fsl[kTileMaxFibers];   // first segment lane of each fiber (packed layout)
    int maxS;
    double bc[5];
    double wr[5 * TpCfg<NT>::waves];
};

// the thread's share of a row: entries/columns [b, e) of a range of n,
// the first half to h = 0, the second to h = 1 (all of it when not split)
__device__ __forceinline__ void share(uint32_t b0, uint32_t b1, bool split, int h, uint32_t& b, uint32_t& e)
{
    if (!split) { b = b0; e = b1; return; }
    const uint32_t mid = b0 + (b1 - b0 + 1) / 2;
    b = h ? mid : b0;
    e = h ? b1 : mid;
}

// Coupled records of a row share: the record (Rec) at col*mul + add for the
// first PRE entries of [e, e1); the rest is gathered when consumed.  One
// record holds every vector the operand needs at that unknown, so a coupled
// point costs one cache line rather than one per vector.
template <int PRE, typename Rec>
struct Gath {
    uint32_t e, e1;
    uint32_t col[PRE];
    double val[PRE];
    Rec g[PRE];
};

// The first PRE entries of the share: addresses clamped to the share's last
// entry and the weights beyond it zero, so the column and value loads are all
// issued before any is used (a load under a per-entry branch made the compiler
// wait for each in turn: one memory round trip per entry).
template <int PRE, typename Rec>
__device__ __forceinline__ void gath_entries(Gath<PRE, Rec>& G, const uint32_t* __restrict__ col,
                                             const double* __restrict__ val)
{
#pragma unroll
    for (int v = 0; v < PRE; ++v) {
        G.col[v] = 0;
        G.val[v] = 0.0;
    }
    if (G.e < G.e1) {
        const uint32_t el = G.e1 - 1;
#pragma unroll
        for (int v = 0; v < PRE; ++v) {
            const uint32_t q = G.e + v < el ? G.e + v : el;
            G.col[v] = col[q];
            const double w = val[q];
            G.val[v] = G.e + v < G.e1 ? w : 0.0;
        }
    }
}

// the records of the prefetched entries (clamped entries repeat the last one)
template <int PRE, typename Rec>
__device__ __forceinline__ void gath_values(Gath<PRE, Rec>& G, const Rec* __restrict__ rec, uint32_t mul,
                                            uint32_t add)
{
    if (G.e < G.e1) {
#pragma unroll
        for (int v = 0; v < PRE; ++v) G.g[v] = rec[G.col[v] * mul + add];
    }
}

// sum of val * X(record) over the share, in entry order
template <int PRE, typename Rec, typename Comb>
__device__ __forceinline__ double gath_sum(const Gath<PRE, Rec>& G, const Rec* __restrict__ rec,
                                           const uint32_t* __restrict__ col, const double* __restrict__ val,
                                           uint32_t mul, uint32_t add, const Comb& X)
{
    double acc = 0.0;
#pragma unroll
    for (int v = 0; v < PRE; ++v)
        if (G.e + v < G.e1) acc += G.val[v] * X(G.g[v]);
    // entries beyond the prefetched ones, a few at a time: the column and value
    // loads of a batch, then its record loads, are issued together (addresses
    // clamped to the share's last entry, weights of the clamped ones zero),
    // and the next batch's columns and values are requested before this
    // batch's records are used, so a long row costs one memory round trip per
    // batch instead of two (or one per entry)
    constexpr int CH = sizeof(Rec) > 16 ? 2 : 4;   // registers of the batch's records
    const uint32_t el = G.e1 - 1;
    uint32_t e = G.e + PRE;
    if (e >= G.e1) return acc;
    uint32_t c[CH];
    double w[CH];
    const auto entries = [&](uint32_t eb, uint32_t* cc, double* ww) {
#pragma unroll
        for (int v = 0; v < CH; ++v) {
            const uint32_t q = eb + v < el ? eb + v : el;
            cc[v] = col[q];
            const double x = val[q];
            ww[v] = eb + v < G.e1 ? x : 0.0;
        }
    };
    entries(e, c, w);
    while (true) {
        Rec r[CH];
#pragma unroll
        for (int v = 0; v < CH; ++v) r[v] = rec[c[v] * mul + add];
        const uint32_t en = e + CH;
        const bool more = en < G.e1;
        uint32_t cn[CH];
        double wn[CH];
        if (more) entries(en, cn, wn);
#pragma unroll
        for (int v = 0; v < CH; ++v) acc += w[v] * X(r[v]);
        if (!more) break;
        e = en;
#pragma unroll
        for (int v = 0; v < CH; ++v) {
            c[v] = cn[v];
            w[v] = wn[v];
        }
    }
    return acc;
}

// phat from the record (shat, w, phat_old, z_old), or from (y, z) in the first
// launch: the expression the owner of the entry evaluates for its own row
struct PhatComb {
    double beta, omega;
    __device__ __forceinline__ double operator()(const double4& r) const
    {
        const double yj = r.x + (-omega) * r.y;
        return beta == 0.0 ? yj : yj + beta * (r.z + (-omega) * r.w);
    }
};
struct PhatFirstComb {
    double beta;
    __device__ __forceinline__ double operator()(const double2& r) const
    {
        return beta == 0.0 ? r.x : r.x + beta * 0.0;
    }
};
// shat from (y, z)
struct ShatComb {
    double alpha;
    __device__ __forceinline__ double operator()(const double2& r) const { return r.x + (-alpha) * r.y; }
};

// Row state of one thread for the operator and the block apply.
template <int D, typename TB, int NB>
struct TpRow {
    int u;             // row in the tile
    bool act, own, split;
    int h;
    // the row's fiber in the tile: index, first point, points; point in the fiber
    int k, fo, np, p;
    // constants of the owner
    double pd0, pd1, df0, df1;
    uint32_t pt, f;
    // block inverse share: columns [j0, j1); in staged launches [j0, jg) come
    // from the LDS stage, where the row's column 0 sits at soff
    uint32_t j0, j1, jg, soff;
    TB blk[NB];
    // segment constants (thread s < ts), masked at the fiber ends after the tables
    double sdf[D], sjd, sja, sjc;
    int sk, sp;
    // packed layout with precomputed reduction coefficients: the segment of the
    // thread's lane (-1: none), its alpha and gamma per level, reduced diagonal
    int psg;
    double pal[kPcrLevels], pga[kPcrLevels], pbf;
};

// the thread's segment (s = threadIdx.x < ts): rfDiff and the JJ' entries of
// global segment S0 + s and its predecessor, loaded before the tables are known
template <int D, typename TB, int NB>
__device__ __forceinline__ void tp_seg_constants(const FiberArgs& a, const Tile& T, TpRow<D, TB, NB>& R)
{
#pragma unroll
    for (int d = 0; d < D; ++d) R.sdf[d] = 0.0;
    R.sjd = 1.0;
    R.sja = R.sjc = 0.0;
    const int s = threadIdx.x;
    if (s < (int)T.ts) {
        const uint32_t g = T.S0 + s;
#pragma unroll
        for (int d = 0; d < D; ++d) R.sdf[d] = a.diff[D * g + d];
        R.sjd = a.jd0[g];
        R.sjc = a.je0[g];
        if (g >= 1) R.sja = a.je0[g - 1];
    }
}

template <int D, int NT>
__device__ __forceinline__ void tp_roles(const Tile& T, int& u, bool& act, bool& own, bool& split, int& h)
{
    constexpr int SP = TpCfg<NT>::split;
    split = T.nu <= (uint32_t)SP;
    const int tid = threadIdx.x;
    u = split ? (tid & (SP - 1)) : tid;
    h = split ? (tid / SP) : 0;
    act = u < (int)T.nu && tid < TpCfg<NT>::work;
    own = act && h == 0;
}

// The row's mB and mC extents, all four loaded before any is used (the
// shares are computed after both loads), then the thread's share of each.
template <int D, typename GBT, typename GCT>
__device__ __forceinline__ void tp_extents(const FiberArgs& a, const Tile& T, int u, bool act, bool split, int h,
                                           uint32_t g, GBT& GB, GCT& GC)
{
    uint32_t b0 = 0, b1 = 0, c0 = 0, c1 = 0;
    if (act) {
        if (a.Buse) {
            const uint32_t r = T.P0 + (uint32_t)u / D;
            b0 = a.Brow[r];
            b1 = a.Brow[r + 1];
        }
        if (a.Cuse) {
            c0 = a.Crow[g];
            c1 = a.Crow[g + 1];
        }
    }
    share(b0, b1, split, h, GB.e, GB.e1);
    share(c0, c1, split, h, GC.e, GC.e1);
}

// P' and rfDiff of the segments after and before the owner's point, by
// global index (masked later with the fiber's extent)
template <int D, typename TB, int NB>
__device__ __forceinline__ void tp_row_constants(const FiberArgs& a, const Tile& T, TpRow<D, TB, NB>& R)
{
    R.pd0 = R.pd1 = R.df0 = R.df1 = 0.0;
    R.pt = 0;
    R.f = 0;
    if (!R.own) return;
    R.pt = (uint32_t)R.u / D;
    const int d = R.u - D * (int)R.pt;
    const uint32_t P = T.P0 + R.pt;
    R.f = a.pfib[P];
    const uint32_t g = P - R.f;     // segment starting at P, if P is not the fiber's last point
    if (g < a.nseg) {
        R.pd0 = a.pdiff[g];
        R.df0 = a.diff[D * g + d];
    }
    if (g >= 1) {
        R.pd1 = a.pdiff[g - 1];
        R.df1 = a.diff[D * (g - 1) + d];
    }
}

// after the tables: the row's fiber, its block-inverse share, and the
// segment's fiber (masking the JJ' couplings at the fiber ends)
template <int D, typename TB, int NB, bool SG = false, typename LDS>
__device__ __forceinline__ void tp_after_tables(const FiberArgs& a, const LDS& L, const Tile& T, const TB* B,
                                                TpRow<D, TB, NB>& R)
{
    R.j0 = R.j1 = R.jg = R.soff = 0;
    R.k = R.fo = R.np = R.p = 0;
    if (R.act) {
        const uint32_t pt = (uint32_t)R.u / D;
        R.k = tile_fiber_of(L.fo, (int)T.nft, pt);
        R.fo = (int)L.fo[R.k];
        R.np = (int)L.fo[R.k + 1] - R.fo;
        R.p = (int)pt - R.fo;
        if (B != nullptr && L.fuse[R.k]) share(0, (uint32_t)(D * R.np), R.split, R.h, R.j0, R.j1);
        R.jg = R.j0;
        if constexpr (SG) {
            // the stage holds the tile's blocks from the 16-byte boundary at or
            // below the first one, up to kStageBytes: whole columns of the
            // row's block that lie inside it are read from LDS
            constexpr uint64_t al = 16 / sizeof(TB), cap = kStageBytes / sizeof(TB);
            const uint64_t e0 = L.fblk[0] - L.fblk[0] % al;
            const uint64_t off = L.fblk[R.k] - e0;
            const uint32_t bs = (uint32_t)(D * R.np);
            const uint32_t jst = off >= cap ? 0u : (uint32_t)((cap - off) / bs < bs ? (cap - off) / bs : bs);
            R.jg = jst < R.j0 ? R.j0 : jst > R.j1 ? R.j1 : jst;
            R.soff = (uint32_t)off + (uint32_t)(R.u - D * R.fo);
        }
    }
    R.sk = R.sp = 0;
    const int s = threadIdx.x;
    if (s < (int)T.ts) {
        const int ks = tile_fiber_of_seg(L.fo, (int)T.nft, (uint32_t)s);
        const int p = s - ((int)L.fo[ks] - ks);
        const int S = (int)L.fo[ks + 1] - (int)L.fo[ks] - 1;
        if (p == 0) R.sja = 0.0;
        if (p + 1 >= S) R.sjc = 0.0;
        R.sk = ks;
        R.sp = p;
    }
}

// The first NB columns of the row's share of the block inverse, requested
// after the coupling gathers of the launch: loads return in order, so a gather
// issued after them would wait for the whole block stream before the operator
// could start; issued here, the stream overlaps the operator's LDS phases.
// (Staged launches: the first NB columns after the staged ones.)
template <int D, typename TB, int NB, typename LDS>
__device__ __forceinline__ void tp_block_prefetch(const LDS& L, const TB* B, TpRow<D, TB, NB>& R)
{
#pragma unroll
    for (int v = 0; v < NB; ++v) R.blk[v] = TB(0);
    const uint32_t jb = R.jg;
    if (R.j1 > jb) {
        const int bs = D * R.np;
        const TB* col = B + L.fblk[R.k] + (R.u - D * R.fo);
        if constexpr (std::is_same<TB, float>::value) {
            // FP32: addresses clamped to the share's last column, entries
            // beyond it zeroed by a select, so the loads are issued together
            const uint32_t jl = R.j1 - 1;
#pragma unroll
            for (int v = 0; v < NB; ++v) {
                const uint32_t j = jb + v < jl ? jb + v : jl;
                const TB e = col[(size_t)bs * j];
                R.blk[v] = jb + v < R.j1 ? e : TB(0);
            }
        } else {
            // FP64: predicated (the branch-free form holds twice the
            // registers live and measures slower)
#pragma unroll
            for (int v = 0; v < NB; ++v)
                if (jb + v < R.j1) R.blk[v] = col[(size_t)bs * (jb + v)];
        }
    }
}

// The staging wavefronts of a staged launch: the tile's block inverses, from
// the 16-byte boundary at or below the first block, copied into the stage by
// global-to-LDS loads of 16 bytes per lane (global_load_lds_dwordx4: M0 holds
// the wavefront's LDS address, lane l writes 16 bytes at M0 + 16 l), each
// wavefront every kTpStageWaves-th KB, up to kStageBytes.  A copy may read up
// to 15 bytes past the tile's last block, inside the buffer's spare elements.
// Nothing is copied for a system that has converged.  The loads are issued
// from inline assembly so that the compiler does not make the workgroup's
// barriers wait for them (a fence after a global-to-LDS load waits for it):
// the copy runs through the prologue, the gathers and the operator, and the
// staging wavefronts wait for it themselves (tp_stage_drain) before the
// barrier ahead of the block apply and before any return.  They issue it last
// before the first barrier, on a path of their own, so that no wait of the
// prologue's (all loads count, in order) stands between it and that barrier.
// Loads the compiler does not count only make its own waits stricter (loads
// complete in order); the loads write no register, and the one wait state
// between the M0 write and the load that reads it is an s_nop.
template <typename TB>
__device__ __forceinline__ void tp_stage_blocks(const KArgs& k, const TB* B, const Tile& T, unsigned char* stage)
{
    const uint64_t b0 = k.blkoff[T.fb], b1 = k.blkoff[T.fb + T.nft];
    if (k.st_in[T.sys].done) return;
    constexpr uint64_t al = 16 / sizeof(TB);
    const uint64_t e0 = b0 - b0 % al;
    const uint64_t all = (b1 - e0) * sizeof(TB);
    const uint32_t nb = (uint32_t)(all < kStageBytes ? all : kStageBytes);
    const unsigned char* src = reinterpret_cast<const unsigned char*>(B + e0);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t w = (threadIdx.x - kTpThreads) / kWave;
    const uint32_t base = (uint32_t)(size_t)(__attribute__((address_space(3))) unsigned char*)stage;
    for (uint32_t o = 16 * kWave * w; o < nb; o += 16 * kWave * kTpStageWaves) {
        const uint32_t lo = o + 16 * lane;
        if (lo < nb) {
            const uint32_t m0 = __builtin_amdgcn_readfirstlane(base + o);
            asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                         :: "s"(m0), "v"(src + lo) : "memory", "m0");
        }
    }
}

// The staging wavefronts wait for their copies (before the barrier that
// precedes any read of the stage, and before they end).
template <bool SG>
__device__ __forceinline__ void tp_stage_drain()
{
    if constexpr (SG)
        if (threadIdx.x >= kTpThreads) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// The packed layout's lane of this thread: the segment it solves and the
// precomputed reduction coefficients of that segment, requested right after
// the tables so that they arrive during the gathers.  Levels past the tile's
// longest fiber are not loaded (the solve stops before them).
template <int D, typename TB, int NB, typename LDS>
__device__ __forceinline__ void tp_pcr_prefetch(const FiberArgs& a, const LDS& L, const Tile& T,
                                                TpRow<D, TB, NB>& R)
{
    R.psg = -1;
#pragma unroll
    for (int l = 0; l < kPcrLevels; ++l) R.pal[l] = R.pga[l] = 0.0;
    R.pbf = 1.0;
    const int s = threadIdx.x;
    if (!T.ws || s >= (int)T.ws) return;
    int lo = 0, hi = (int)T.nft;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (L.fsl[mid] <= (uint32_t)s) lo = mid;
        else hi = mid;
    }
    const int off = s - (int)L.fsl[lo];
    if (off >= (int)(L.fo[lo + 1] - L.fo[lo]) - 1) return;
    R.psg = (int)L.fo[lo] - lo + off;
    const size_t g = T.S0 + (uint32_t)R.psg, ns = a.nseg;
#pragma unroll
    for (int l = 0; l < kPcrLevels; ++l) {
        if ((1 << l) < L.maxS) {
            R.pal[l] = a.pcr[(size_t)(2 * l) * ns + g];
            R.pga[l] = a.pcr[(size_t)(2 * l + 1) * ns + g];
        }
    }
    R.pbf = a.pcr[(size_t)(2 * kPcrLevels) * ns + g];
}

// The operator over the tile once L.x holds the operand and the helpers'
// coupling sums are in L.hb/L.hc: forces, multipliers, L.y.  bsum/csum: the
// owner's own shares.  Ends synchronised.
template <int D, typename TB, int NB, bool PK, bool SG = false, typename LDS>
__device__ __forceinline__ void tp_operator(LDS& L, const Tile& T, const TpRow<D, TB, NB>& R, double bsum,
                                            double csum)
{
    if (R.own) {
        const int k = R.k;
        const int fo = R.fo;
        const int np = R.np;
        const int p = R.p;
        const int d = R.u - D * (int)R.pt;
        const int S = np - 1;
        double fp = bsum;
        double cs = csum;
        if (R.split) {
            fp = fp + L.hb[R.u];
            cs = cs + L.hc[R.u];
        }
        fp += cs;
        const double* xf = L.x + D * fo;
        if (np > 2) {
            const double kr = L.fk[k];
            double acc = 0.0;
            if (p <= np - 3) {
                const double v = kr * (xf[D * p + d] - 2 * xf[D * (p + 1) + d] + xf[D * (p + 2) + d]);
                acc -= v;
            }
            if (p >= 1 && p <= np - 2) {
                const double v = kr * (xf[D * (p - 1) + d] - 2 * xf[D * p + d] + xf[D * (p + 1) + d]);
                acc += v + v;
            }
            if (p >= 2) {
                const double v = kr * (xf[D * (p - 2) + d] - 2 * xf[D * (p - 1) + d] + xf[D * p + d]);
                acc -= v;
            }
            fp += acc;
        }
        if (p < S && R.pd0 != 0.0) fp += R.pd0 * (xf[D * (p + 1) + d] - xf[D * p + d]);
        if (p >= 1 && R.pd1 != 0.0) fp -= R.pd1 * (xf[D * p + d] - xf[D * (p - 1) + d]);
        L.f[R.u] = fp;
    }
    __syncthreads();
    const int ts = (int)T.ts;
    const int s = threadIdx.x;
    double t = 0.0;
    if (s < ts) {
        const int pt = (int)L.fo[R.sk] + R.sp;
#pragma unroll
        for (int d = 0; d < D; ++d) t += R.sdf[d] * (L.f[D * (pt + 1) + d] - L.f[D * pt + d]);
    }
    if (PK && T.ws && T.ws <= blockDim.x) {
        // packed layout with precomputed coefficients: only the right-hand
        // sides move, each level two shuffles and two products
        if (s < ts) L.lam[s] = t;
        __syncthreads();
        if ((s & ~(kWave - 1)) < (int)T.ws) {
            const int lane = s & (kWave - 1);
            double dv = R.psg >= 0 ? L.lam[R.psg] : 0.0;
#pragma unroll
            for (int l = 0; l < kPcrLevels; ++l) {
                const int k = 1 << l;
                if (k >= L.maxS) break;
                double dl = __shfl_up(dv, k, kWave), du = __shfl_down(dv, k, kWave);
                if (lane < k) dl = 0.0;
                if (lane + k >= kWave) du = 0.0;
                dv = dv + R.pal[l] * dl + R.pga[l] * du;
            }
            const double lam = dv / R.pbf;
            if (R.psg >= 0) L.lam[R.psg] = lam;
        }
    } else if (T.ws && T.ws <= blockDim.x) {
        // packed layout: the segment rows are handed through LDS to the lanes
        // of the packed layout, where every wavefront holds whole fibers and
        // solves them with shuffles (one barrier instead of two per step)
        if (s < ts) {
            L.pa[s] = R.sja;
            L.pb[s] = R.sjd;
            L.pc[s] = R.sjc;
            L.lam[s] = t;
        }
        __syncthreads();
        if ((s & ~(kWave - 1)) < (int)T.ws) {
            double ra = 0.0, rb = 1.0, rc = 0.0, rd = 0.0;
            int sg = -1;
            if (s < (int)T.ws) {
                int lo = 0, hi = (int)T.nft;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (L.fsl[mid] <= (uint32_t)s) lo = mid;
                    else hi = mid;
                }
                const int off = s - (int)L.fsl[lo];
                if (off < (int)(L.fo[lo + 1] - L.fo[lo]) - 1) {
                    sg = (int)L.fo[lo] - lo + off;
                    ra = L.pa[sg];
                    rb = L.pb[sg];
                    rc = L.pc[sg];
                    rd = L.lam[sg];
                }
            }
            const double lam = pcr_solve_span(kWave, L.maxS, ra, rb, rc, rd);
            if (sg >= 0) L.lam[sg] = lam;
        }
    } else if (ts <= kWave) {
        if (s < kWave) {
            const double lam = pcr_solve_span(ts, L.maxS, R.sja, R.sjd, R.sjc, t);
            if (s < ts) L.lam[s] = lam;
        }
    } else {
        const double lam = pcr_lds(L.pa, L.pb, L.pc, L.lam, ts, L.maxS, R.sja, R.sjd, R.sjc, t);
        __syncthreads();
        if (s < ts) L.lam[s] = lam;
    }
    __syncthreads();
    if (R.own) {
        const int k = R.k;
        const int fo = R.fo;
        const int S = R.np - 1;
        const int p = R.p;
        const int sl = fo - k + p;
        double v = L.f[R.u];
        if (p < S) v = v + R.df0 * L.lam[sl];
        if (p >= 1) v = v - R.df1 * L.lam[sl - 1];
        const double yv = L.fsca[k] * v;
        L.y[R.u] = yv + L.x[R.u];
    }
    tp_stage_drain<SG>();   // the barrier before the block apply reads the stage
    __syncthreads();
}

// out[u] = Binv xin over the tile: each thread its column share, the helper's
// partial through LDS; copies where there is no usable block.  Ends synchronised.
template <int D, typename TB, int NB, int CH = kApplyChunk, bool SG = false, typename LDS>
__device__ __forceinline__ double tp_block_apply(LDS& L, const Tile& T, const TpRow<D, TB, NB>& R, const TB* B,
                                                 const double* xin, const TB* stage = nullptr)
{
    double acc[4] = { 0, 0, 0, 0 };
    const int k = R.k, fo = R.fo, bs = D * R.np;
    const bool use = R.act && B != nullptr && L.fuse[k];
    uint32_t rot = 0;
    if (use && R.j1 > R.j0) {
        const double* xs = xin + D * fo;
        if constexpr (SG) {
            // staged columns [j0, jg) from LDS; column j accumulates in slot
            // (j - j0) & 3 as in the other launches, so that the result is the
            // same bit for bit: the slots are rotated for the columns after jg
            // and rotated back at the end
            const TB* sl = stage + R.soff;
            uint32_t j = R.j0;
#pragma unroll 2
            for (; j + 4 <= R.jg; j += 4) {
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[v] += (double)sl[bs * (j + v)] * xs[j + v];
            }
            rot = R.jg - j;
            if (rot > 0) acc[0] += (double)sl[bs * j] * xs[j];
            if (rot > 1) acc[1] += (double)sl[bs * (j + 1)] * xs[j + 1];
            if (rot > 2) acc[2] += (double)sl[bs * (j + 2)] * xs[j + 2];
            const double a0 = acc[0], a1 = acc[1], a2 = acc[2], a3 = acc[3];
            acc[0] = rot == 0 ? a0 : rot == 1 ? a1 : rot == 2 ? a2 : a3;
            acc[1] = rot == 0 ? a1 : rot == 1 ? a2 : rot == 2 ? a3 : a0;
            acc[2] = rot == 0 ? a2 : rot == 1 ? a3 : rot == 2 ? a0 : a1;
            acc[3] = rot == 0 ? a3 : rot == 1 ? a0 : rot == 2 ? a1 : a2;
        }
        // Branch-free: every operand read stays inside the
        // fiber's rows (the index is clamped to the last column of this
        // thread's share) and is multiplied by a zero entry beyond it, so the
        // LDS reads of the loop are in flight together instead of one per
        // branch (16.1 instead of 17.5 us per launch on the 1000-fiber system).
        const uint32_t jb = R.jg;
        const uint32_t jl = R.j1 - 1;
#pragma unroll
        for (int v = 0; v < NB; ++v) {
            const uint32_t j = jb + v < jl ? jb + v : jl;
            acc[v & 3] += (double)R.blk[v] * xs[j];
        }
        const TB* col = B + L.fblk[k] + (R.u - D * fo);
        for (uint32_t j = jb + NB; j < R.j1; j += CH) {
            TB c[CH];
            if constexpr (std::is_same<TB, float>::value) {
                // FP32: clamped addresses and a select (no branch per load), the
                // chunk's loads in flight together (a predicated load per column
                // cost FP32 blocks 40-60% of the launch on tiles of one thread
                // per row)
#pragma unroll
                for (int v = 0; v < CH; ++v) {
                    const uint32_t jj = j + v < jl ? j + v : jl;
                    const TB e = col[(size_t)bs * jj];
                    c[v] = j + v < R.j1 ? e : TB(0);
                }
            } else {
#pragma unroll
                for (int v = 0; v < CH; ++v)
                    c[v] = (j + v < R.j1) ? col[(size_t)bs * (j + v)] : TB(0);
            }
#pragma unroll
            for (int v = 0; v < CH; ++v) {
                const uint32_t jj = j + v < jl ? j + v : jl;
                acc[v & 3] += (double)c[v] * xs[jj];
            }
        }
    }
    if constexpr (SG) {
        const double a0 = acc[0], a1 = acc[1], a2 = acc[2], a3 = acc[3];
        acc[0] = rot == 0 ? a0 : rot == 1 ? a3 : rot == 2 ? a2 : a1;
        acc[1] = rot == 0 ? a1 : rot == 1 ? a0 : rot == 2 ? a3 : a2;
        acc[2] = rot == 0 ? a2 : rot == 1 ? a1 : rot == 2 ? a0 : a3;
        acc[3] = rot == 0 ? a3 : rot == 1 ? a2 : rot == 2 ? a1 : a0;
    }
    const double part = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    if (R.split && R.act && R.h == 1) L.hb[R.u] = part;
    __syncthreads();
    if (!R.own) return 0.0;
    if (!use) return xin[R.u];
    return R.split ? part + L.hb[R.u] : part;
}

// Partial sums of the previous launch re-reduced by wavefront 1 of every tile
// of the system (same order everywhere, so every tile derives the same
// scalars), and the scalar updates of the P and Q launches.
template <bool FIRST>
__device__ __forceinline__ void tp_partials(const KArgs& k, const Tile& T, double* bc)
{
    const uint32_t nt = k.ta.ntiles;
    if (threadIdx.x >= kWave && threadIdx.x < 2 * kWave) {
        if (FIRST) {
            const double rr = wsum(k.part + P_RR * nt, T.tb, T.te);
            if (threadIdx.x == kWave) bc[0] = rr;
        } else {
            double S4[4];
            wsum4(k.part4, T.tb, T.te, S4);
            const double m = wmax(k.part + P_SMAX * nt, T.tb, T.te);
            if (threadIdx.x == kWave) { bc[0] = S4[0]; bc[1] = S4[1]; bc[2] = S4[2]; bc[3] = S4[3]; bc[4] = m; }
        }
    }
}

__device__ __forceinline__ void tp_scalars(SysScalars& sc, const double* bc)
{
    sc.pending = 0;
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
}

__device__ __forceinline__ void tq_partials(const KArgs& k, const Tile& T, double* bc)
{
    const uint32_t nt = k.ta.ntiles;
    if (threadIdx.x >= kWave && threadIdx.x < 2 * kWave) {
        const double res = wmax(k.part + P_RMAX * nt, T.tb, T.te);
        const double dsum = wsum(k.part + P_DELTA * nt, T.tb, T.te);
        if (threadIdx.x == kWave) { bc[0] = res; bc[1] = dsum; }
    }
}

template <int PC>
__device__ __forceinline__ void tq_scalars(SysScalars& sc, const double* bc)
{
    decide(sc, bc[0]);
    if (!sc.done) {
        const double delta = bc[1];
        if (PC != CYMG_PRECOND_NONE && delta == 0.0) { sc.done = 1; sc.flag = 4; }
        else sc.alpha = sc.rho1 / delta;
    }
    sc.stage = ST_C;
}

// sums of the wavefronts' partials in wavefront order (NaN-propagating max)
__device__ __forceinline__ double waves_max(const double* w, int nw)
{
    double m = w[0];
    for (int q = 1; q < nw; ++q) m = nan_max(m, w[q]);
    return m;
}
__device__ __forceinline__ double waves_sum(const double* w, int nw)
{
    double t = w[0];
    for (int q = 1; q < nw; ++q) t += w[q];
    return t;
}

// Solid tiles of workgroups smaller than kTpThreads: the same arithmetic in
// the same order, each row's vectors read inside its loop rather than held
// for all of the thread's rows from the start (which would take the launch
// past 128 VGPRs and a quarter of its resident workgroups on fiber tiles).
template <int D, int PC, bool FIRST, int NT>
__device__ void solid_tile_p_lean(const KArgs& k, const Tile& T, SolidTileLds& S)
{
    const uint32_t ks = T.fb - k.nfiber, U0 = T.U0, nu = T.nu, nt = k.ta.ntiles;
    const int nw = blockDim.x / kWave;
    SysScalars sc = state_load(k.st_in + T.sys);
    tp_partials<FIRST>(k, T, S.bc);
    __syncthreads();
    if (sc.done) return;
    tp_scalars(sc, S.bc);
    if (sc.done) return;
    double rmax = 0.0, dl = 0.0;
    for (uint32_t u = threadIdx.x; u < nu; u += NT) {
        const uint32_t g = U0 + u;
        const double4 q1 = k.g1a[g];
        double ri = k.r[g];
        double yi = FIRST ? k.g2[g].x : 0.0;
        if (sc.pending) {
            k.x[g] = k.x[g] + sc.omega * q1.x;
            ri = ri + (-sc.omega) * k.t[g];
            k.r[g] = ri;
            yi = q1.x + (-sc.omega) * q1.y;
        }
        rmax = nan_max(rmax, fabs(ri));
        S.a[u] = sc.beta == 0.0 ? yi : yi + sc.beta * (q1.z + (-sc.omega) * q1.w);
        S.e[u] = yi;
    }
    __syncthreads();
    const auto O = [&](uint32_t i, double v) { S.y[i - U0] = v; };
    if constexpr (FIRST) {
        const PhatFirstComb C{ sc.beta };
        const auto X = [&](uint32_t i) -> double { return i - U0 < nu ? S.a[i - U0] : C(k.g2[i]); };
        solid_operator<D>(k.fa, k.sa, ks, X, O, S.red, S.lrow, kSolidTileRows + 1);
    } else {
        const PhatComb C{ sc.beta, sc.omega };
        const auto X = [&](uint32_t i) -> double { return i - U0 < nu ? S.a[i - U0] : C(k.g1a[i]); };
        solid_operator<D>(k.fa, k.sa, ks, X, O, S.red, S.lrow, kSolidTileRows + 1);
    }
    __syncthreads();
    for (uint32_t u = threadIdx.x; u < nu; u += NT) {
        const double v = S.y[u];
        k.v[U0 + u] = v;
        dl += k.rt[U0 + u] * v;
    }
    rmax = wave_max(rmax);
    dl = wave_sum(dl);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        S.wr[threadIdx.x / kWave] = rmax;
        S.wr[16 + threadIdx.x / kWave] = dl;
    }
    const auto Rec = [&](uint32_t i, double z) {
        const uint32_t u = i - U0;
        k.g2[i] = make_double2(S.e[u], z);
        reinterpret_cast<double2*>(k.g1b + i)[1] = make_double2(S.a[u], z);
    };
    if (PC == CYMG_PRECOND_NONE) {
        for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x) Rec(U0 + u, S.y[u]);
        __syncthreads();
    } else {
        const auto In = [&](uint32_t i) -> double { return S.y[i - U0]; };
        solid_precond<D>(k.sa, ks, In, Rec, S.red);
    }
    if (threadIdx.x == 0) {
        k.part[P_RMAX * nt + T.t] = waves_max(S.wr, nw);
        k.part[P_DELTA * nt + T.t] = waves_sum(S.wr + 16, nw);
    }
}

template <int D, int PC, int NT>
__device__ void solid_tile_q_lean(const KArgs& k, const Tile& T, SolidTileLds& S)
{
    const uint32_t ks = T.fb - k.nfiber, U0 = T.U0, nu = T.nu, nt = k.ta.ntiles;
    const int nw = blockDim.x / kWave;
    SysScalars sc = state_load(k.st_in + T.sys);
    tq_partials(k, T, S.bc);
    __syncthreads();
    if (sc.done) return;
    tq_scalars<PC>(sc, S.bc);
    if (sc.done) return;
    double smax = 0.0;
    for (uint32_t u = threadIdx.x; u < nu; u += NT) {
        const uint32_t g = U0 + u;
        const double si = k.r[g] + (-sc.alpha) * k.v[g];
        k.r[g] = si;
        k.x[g] = k.x[g] + sc.alpha * k.g1b[g].z;
        smax = nan_max(smax, fabs(si));
        const double2 q2 = k.g2[g];
        S.a[u] = q2.x + (-sc.alpha) * q2.y;
    }
    __syncthreads();
    {
        const ShatComb C{ sc.alpha };
        const auto X = [&](uint32_t i) -> double { return i - U0 < nu ? S.a[i - U0] : C(k.g2[i]); };
        const auto O = [&](uint32_t i, double v) { S.y[i - U0] = v; };
        solid_operator<D>(k.fa, k.sa, ks, X, O, S.red, S.lrow, kSolidTileRows + 1);
    }
    __syncthreads();
    double tt = 0, ts = 0, rts = 0, rtt = 0;
    for (uint32_t u = threadIdx.x; u < nu; u += NT) {
        const double ti = S.y[u];
        const double si = k.r[U0 + u];
        const double vrt = k.rt[U0 + u];
        k.t[U0 + u] = ti;
        tt += ti * ti;
        ts += ti * si;
        rts += vrt * si;
        rtt += vrt * ti;
    }
    smax = wave_max(smax);
    tt = wave_sum(tt);
    ts = wave_sum(ts);
    rts = wave_sum(rts);
    rtt = wave_sum(rtt);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const int w = threadIdx.x / kWave;
        S.wr[w] = smax;
        S.wr[16 + w] = tt;
        S.wr[32 + w] = ts;
        S.wr[48 + w] = rts;
        S.wr[64 + w] = rtt;
    }
    const auto Rec = [&](uint32_t i, double w) {
        reinterpret_cast<double2*>(k.g1b + i)[0] = make_double2(S.a[i - U0], w);
    };
    if (PC == CYMG_PRECOND_NONE) {
        for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x) Rec(U0 + u, S.y[u]);
        __syncthreads();
    } else {
        const auto In = [&](uint32_t i) -> double { return S.y[i - U0]; };
        solid_precond<D>(k.sa, ks, In, Rec, S.red);
    }
    if (threadIdx.x == 0) {
        k.part[P_SMAX * nt + T.t] = waves_max(S.wr, nw);
        *reinterpret_cast<double4*>(k.part4 + 4 * (size_t)T.t) =
            make_double4(waves_sum(S.wr + 16, nw), waves_sum(S.wr + 32, nw), waves_sum(S.wr + 48, nw),
                         waves_sum(S.wr + 64, nw));
    }
}

// k_tp for a Solid tile (NT threads, kSolidTileRows / NT rows each)
template <int D, int PC, bool FIRST, int NT>
__device__ void solid_tile_p(const KArgs& k, const Tile& T, SolidTileLds& S)
{
    if constexpr (NT < kTpThreads) {
        solid_tile_p_lean<D, PC, FIRST, NT>(k, T, S);
        return;
    }
    constexpr int RS = kSolidTileRows / NT;
    const uint32_t ks = T.fb - k.nfiber, U0 = T.U0, nu = T.nu, nt = k.ta.ntiles;
    const int nw = blockDim.x / kWave;
    double vr[RS], vt[RS], vx[RS], vrt[RS], vy[RS];
    double4 q1[RS];
#pragma unroll
    for (int q = 0; q < RS; ++q) {
        const uint32_t u = threadIdx.x + q * NT;
        vr[q] = vt[q] = vx[q] = vrt[q] = vy[q] = 0.0;
        q1[q] = make_double4(0, 0, 0, 0);
        if (u < nu) {
            const uint32_t g = U0 + u;
            vr[q] = k.r[g]; vt[q] = k.t[g]; vx[q] = k.x[g]; vrt[q] = k.rt[g];
            q1[q] = k.g1a[g];
            if (FIRST) vy[q] = k.g2[g].x;
        }
    }
    SysScalars sc = state_load(k.st_in + T.sys);
    tp_partials<FIRST>(k, T, S.bc);
    __syncthreads();
    if (sc.done) return;
    tp_scalars(sc, S.bc);
    if (sc.done) return;
    double rmax = 0.0, dl = 0.0;
#pragma unroll
    for (int q = 0; q < RS; ++q) {
        const uint32_t u = threadIdx.x + q * NT;
        if (u >= nu) continue;
        const uint32_t g = U0 + u;
        double ri = vr[q], yi = vy[q];
        if (sc.pending) {
            k.x[g] = vx[q] + sc.omega * q1[q].x;
            ri = ri + (-sc.omega) * vt[q];
            k.r[g] = ri;
            yi = q1[q].x + (-sc.omega) * q1[q].y;
        }
        rmax = nan_max(rmax, fabs(ri));
        S.a[u] = sc.beta == 0.0 ? yi : yi + sc.beta * (q1[q].z + (-sc.omega) * q1[q].w);
        S.e[u] = yi;
    }
    __syncthreads();
    const auto O = [&](uint32_t i, double v) { S.y[i - U0] = v; };
    if constexpr (FIRST) {
        const PhatFirstComb C{ sc.beta };
        const auto X = [&](uint32_t i) -> double { return i - U0 < nu ? S.a[i - U0] : C(k.g2[i]); };
        solid_operator<D>(k.fa, k.sa, ks, X, O, S.red, S.lrow, kSolidTileRows + 1);
    } else {
        const PhatComb C{ sc.beta, sc.omega };
        const auto X = [&](uint32_t i) -> double { return i - U0 < nu ? S.a[i - U0] : C(k.g1a[i]); };
        solid_operator<D>(k.fa, k.sa, ks, X, O, S.red, S.lrow, kSolidTileRows + 1);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RS; ++q) {
        const uint32_t u = threadIdx.x + q * NT;
        if (u >= nu) continue;
        const double v = S.y[u];
        k.v[U0 + u] = v;
        dl += vrt[q] * v;
    }
    rmax = wave_max(rmax);
    dl = wave_sum(dl);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        S.wr[threadIdx.x / kWave] = rmax;
        S.wr[16 + threadIdx.x / kWave] = dl;
    }
    // z = M^-1 v; records (y, z) and (phat, z)
    const auto Rec = [&](uint32_t i, double z) {
        const uint32_t u = i - U0;
        k.g2[i] = make_double2(S.e[u], z);
        reinterpret_cast<double2*>(k.g1b + i)[1] = make_double2(S.a[u], z);
    };
    if (PC == CYMG_PRECOND_NONE) {
        for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x) Rec(U0 + u, S.y[u]);
        __syncthreads();
    } else {
        const auto In = [&](uint32_t i) -> double { return S.y[i - U0]; };
        solid_precond<D>(k.sa, ks, In, Rec, S.red);
    }
    if (threadIdx.x == 0) {
        k.part[P_RMAX * nt + T.t] = waves_max(S.wr, nw);
        k.part[P_DELTA * nt + T.t] = waves_sum(S.wr + 16, nw);
    }
}

// k_tq for a Solid tile
template <int D, int PC, int NT>
__device__ void solid_tile_q(const KArgs& k, const Tile& T, SolidTileLds& S)
{
    if constexpr (NT < kTpThreads) {
        solid_tile_q_lean<D, PC, NT>(k, T, S);
        return;
    }
    constexpr int RS = kSolidTileRows / NT;
    const uint32_t ks = T.fb - k.nfiber, U0 = T.U0, nu = T.nu, nt = k.ta.ntiles;
    const int nw = blockDim.x / kWave;
    double vr[RS], vv[RS], vx[RS], vph[RS], vrt[RS], si[RS];
    double2 q2[RS];
#pragma unroll
    for (int q = 0; q < RS; ++q) {
        const uint32_t u = threadIdx.x + q * NT;
        vr[q] = vv[q] = vx[q] = vph[q] = vrt[q] = si[q] = 0.0;
        q2[q] = make_double2(0, 0);
        if (u < nu) {
            const uint32_t g = U0 + u;
            vr[q] = k.r[g]; vv[q] = k.v[g]; vx[q] = k.x[g]; vrt[q] = k.rt[g];
            q2[q] = k.g2[g];
            vph[q] = k.g1b[g].z;
        }
    }
    SysScalars sc = state_load(k.st_in + T.sys);
    tq_partials(k, T, S.bc);
    __syncthreads();
    if (sc.done) return;
    tq_scalars<PC>(sc, S.bc);
    if (sc.done) return;
    double smax = 0.0;
#pragma unroll
    for (int q = 0; q < RS; ++q) {
        const uint32_t u = threadIdx.x + q * NT;
        if (u >= nu) continue;
        const uint32_t g = U0 + u;
        si[q] = vr[q] + (-sc.alpha) * vv[q];
        k.r[g] = si[q];
        k.x[g] = vx[q] + sc.alpha * vph[q];
        smax = nan_max(smax, fabs(si[q]));
        S.a[u] = q2[q].x + (-sc.alpha) * q2[q].y;
    }
    __syncthreads();
    {
        const ShatComb C{ sc.alpha };
        const auto X = [&](uint32_t i) -> double { return i - U0 < nu ? S.a[i - U0] : C(k.g2[i]); };
        const auto O = [&](uint32_t i, double v) { S.y[i - U0] = v; };
        solid_operator<D>(k.fa, k.sa, ks, X, O, S.red, S.lrow, kSolidTileRows + 1);
    }
    __syncthreads();
    double tt = 0, ts = 0, rts = 0, rtt = 0;
#pragma unroll
    for (int q = 0; q < RS; ++q) {
        const uint32_t u = threadIdx.x + q * NT;
        if (u >= nu) continue;
        const double ti = S.y[u];
        k.t[U0 + u] = ti;
        tt += ti * ti;
        ts += ti * si[q];
        rts += vrt[q] * si[q];
        rtt += vrt[q] * ti;
    }
    smax = wave_max(smax);
    tt = wave_sum(tt);
    ts = wave_sum(ts);
    rts = wave_sum(rts);
    rtt = wave_sum(rtt);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const int w = threadIdx.x / kWave;
        S.wr[w] = smax;
        S.wr[16 + w] = tt;
        S.wr[32 + w] = ts;
        S.wr[48 + w] = rts;
        S.wr[64 + w] = rtt;
    }
    // w = M^-1 t; records (shat, w)
    const auto Rec = [&](uint32_t i, double w) {
        reinterpret_cast<double2*>(k.g1b + i)[0] = make_double2(S.a[i - U0], w);
    };
    if (PC == CYMG_PRECOND_NONE) {
        for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x) Rec(U0 + u, S.y[u]);
        __syncthreads();
    } else {
        const auto In = [&](uint32_t i) -> double { return S.y[i - U0]; };
        solid_precond<D>(k.sa, ks, In, Rec, S.red);
    }
    if (threadIdx.x == 0) {
        k.part[P_SMAX * nt + T.t] = waves_max(S.wr, nw);
        *reinterpret_cast<double4*>(k.part4 + 4 * (size_t)T.t) =
            make_double4(waves_sum(S.wr + 16, nw), waves_sum(S.wr + 32, nw), waves_sum(S.wr + 48, nw),
                         waves_sum(S.wr + 64, nw));
    }
}

// at most 128 VGPRs on 256-thread tiles: four resident workgroups per compute unit
template <int NT>
constexpr int kTpWavesPerEu = NT == 256 ? 4 : 1;

// The LDS stage of a staged launch (a separate array, so that LDS accesses to
// the tile tables are known not to touch it and wait for no copy); 16 bytes
// elsewhere.
template <int NT>
constexpr uint32_t kStageOf = NT == kTpStaged ? kStageBytes : 16u;

template <int D, int PC, bool FIRST, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kTpWavesPerEu<NT>))) void k_tp(KArgs k)
{
    using TB = typename PcBlocks<PC>::T;
    using Rec = typename std::conditional<FIRST, double2, double4>::type;
    constexpr int NW = TpCfg<NT>::waves;
    // a workgroup serves a fiber tile or a Solid tile: the two layouts share the LDS
    __shared__ union {
        TpLds<D, NT> t;
        SolidTileLds s;
    } sh;
    TpLds<D, NT>& L = sh.t;
    constexpr bool SG = NT == kTpStaged;
    __shared__ __attribute__((aligned(16))) unsigned char stage[kStageOf<NT>];
    const FiberArgs& a = k.fa;
    const Tile T = tile_begin<D>(k.ta, blockIdx.x);
    if (T.nft == 0) {
        if constexpr (!SG) solid_tile_p<D, PC, FIRST, NT>(k, T, sh.s);
        return;
    }
    TpRow<D, TB, TpCfg<NT>::blk> R;
    tp_roles<D, NT>(T, R.u, R.act, R.own, R.split, R.h);
    // prologue: everything independent of the iteration scalars
    const uint32_t g = T.U0 + R.u;
    double vr = 0, vt = 0, vx = 0, vrt = 0;
    double4 q1 = make_double4(0, 0, 0, 0);
    double vy = 0;
    const Rec* rec;
    if constexpr (FIRST) rec = k.g2;
    else rec = k.g1a;
    Gath<TpCfg<NT>::preb, Rec> GB;
    Gath<TpCfg<NT>::prec, Rec> GC;
    GB.e = GB.e1 = GC.e = GC.e1 = 0;
    const int dd = R.u - D * (R.u / D);
    SysScalars sc;
    if (SG && threadIdx.x >= kTpThreads) {
        // staging wavefronts: the copy, issued last before the first barrier
        // so that no wait of the prologue's falls between the two
        tp_row_constants<D, TB>(a, T, R);
        tp_seg_constants<D, TB>(a, T, R);
        sc = state_load(k.st_in + T.sys);
        if constexpr (SG && PC != CYMG_PRECOND_NONE) tp_stage_blocks<TB>(k, PcBlocks<PC>::get(k), T, stage);
        __syncthreads();
    } else {
        if (R.own) {
            vr = k.r[g]; vt = k.t[g]; vx = k.x[g]; vrt = k.rt[g];
            q1 = k.g1a[g];
            if (FIRST) vy = k.g2[g].x;
        }
        tp_extents<D>(a, T, R.u, R.act, R.split, R.h, g, GB, GC);
        gath_entries(GB, a.Bcol, a.Bval);
        gath_entries(GC, a.Ccol, a.Cval);
        gath_values(GB, rec, (uint32_t)D, (uint32_t)dd);
        gath_values(GC, rec, 1u, 0u);
        tp_row_constants<D, TB>(a, T, R);
        tp_seg_constants<D, TB>(a, T, R);
        sc = state_load(k.st_in + T.sys);
        tile_tables<PC != CYMG_PRECOND_NONE>(a, L, T, k.blkoff, k.useblk);
        if (T.ws && threadIdx.x < T.nft) L.fsl[threadIdx.x] = k.ta.fslot[T.fb + threadIdx.x];
        tp_partials<FIRST>(k, T, L.bc);
        __syncthreads();
    }
    const uint32_t nt = k.ta.ntiles;
    const bool first = T.t == T.tb;
    if (sc.done) {
        if (first && threadIdx.x == 0) state_store(k.st_out + T.sys, sc);
        tp_stage_drain<SG>();
        return;
    }
    tp_scalars(sc, L.bc);
    if (first && threadIdx.x == 0) state_store(k.st_out + T.sys, sc);
    if (sc.done) {
        tp_stage_drain<SG>();
        return;
    }
    tp_after_tables<D, TB, TpCfg<NT>::blk, SG>(a, L, T, PcBlocks<PC>::get(k), R);
    if constexpr (TpCfg<NT>::work == kTpThreads) tp_pcr_prefetch<D, TB>(a, L, T, R);
    // vector updates and phat; q1 = (shat, w, phat_old, z_old) of the row
    double rmax = 0.0, yi = vy, phi = 0.0;
    if (R.own) {
        double ri = vr;
        if (sc.pending) {
            k.x[g] = vx + sc.omega * q1.x;
            ri = ri + (-sc.omega) * vt;
            k.r[g] = ri;
            yi = q1.x + (-sc.omega) * q1.y;
        }
        rmax = fabs(ri);
        phi = sc.beta == 0.0 ? yi : yi + sc.beta * (q1.z + (-sc.omega) * q1.w);
        L.x[R.u] = phi;
    }
    double bsum = 0.0, csum = 0.0;
    if constexpr (FIRST) {
        const PhatFirstComb X{ sc.beta };
        bsum = gath_sum(GB, rec, a.Bcol, a.Bval, (uint32_t)D, (uint32_t)dd, X);
        csum = gath_sum(GC, rec, a.Ccol, a.Cval, 1u, 0u, X);
    } else {
        const PhatComb X{ sc.beta, sc.omega };
        bsum = gath_sum(GB, rec, a.Bcol, a.Bval, (uint32_t)D, (uint32_t)dd, X);
        csum = gath_sum(GC, rec, a.Ccol, a.Cval, 1u, 0u, X);
    }
    if (R.split && R.act && R.h == 1) {
        L.hb[R.u] = bsum;
        L.hc[R.u] = csum;
    }
    tp_block_prefetch<D, TB>(L, PcBlocks<PC>::get(k), R);
    __syncthreads();
    tp_operator<D, TB, TpCfg<NT>::blk, TpCfg<NT>::work == kTpThreads, SG>(L, T, R, bsum, csum);
    double dl = 0.0;
    if (R.own) {
        k.v[g] = L.y[R.u];
        dl = vrt * L.y[R.u];
    }
    rmax = wave_max(rmax);
    dl = wave_sum(dl);
    if ((threadIdx.x & (kWave - 1)) == 0 && threadIdx.x < TpCfg<NT>::work) {
        L.wr[threadIdx.x / kWave] = rmax;
        L.wr[NW + threadIdx.x / kWave] = dl;
    }
    const double zi = tp_block_apply<D, TB, TpCfg<NT>::blk, TpCfg<NT>::chunk, SG>(
        L, T, R, PcBlocks<PC>::get(k), L.y, reinterpret_cast<const TB*>(stage));
    if (R.own) {
        k.g2[g] = make_double2(yi, zi);
        reinterpret_cast<double2*>(k.g1b + g)[1] = make_double2(phi, zi);
    }
    if (threadIdx.x == 0) {
        double m = L.wr[0], d = L.wr[NW];
#pragma unroll
        for (int q = 1; q < NW; ++q) {
            m = nan_max(m, L.wr[q]);
            d += L.wr[NW + q];
        }
        k.part[P_RMAX * nt + T.t] = m;
        k.part[P_DELTA * nt + T.t] = d;
    }
}

template <int D, int PC, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kTpWavesPerEu<NT>))) void k_tq(KArgs k)
{
    using TB = typename PcBlocks<PC>::T;
    constexpr int NW = TpCfg<NT>::waves;
    __shared__ union {
        TpLds<D, NT> t;
        SolidTileLds s;
    } sh;
    TpLds<D, NT>& L = sh.t;
    constexpr bool SG = NT == kTpStaged;
    __shared__ __attribute__((aligned(16))) unsigned char stage[kStageOf<NT>];
    const FiberArgs& a = k.fa;
    const Tile T = tile_begin<D>(k.ta, blockIdx.x);
    if (T.nft == 0) {
        if constexpr (!SG) solid_tile_q<D, PC, NT>(k, T, sh.s);
        return;
    }
    TpRow<D, TB, TpCfg<NT>::blk> R;
    tp_roles<D, NT>(T, R.u, R.act, R.own, R.split, R.h);
    const uint32_t g = T.U0 + R.u;
    double vr = 0, vv = 0, vx = 0, vph = 0, vrt = 0;
    double2 q2 = make_double2(0, 0);
    const double2* rec = k.g2;
    Gath<TpCfg<NT>::preb, double2> GB;
    Gath<TpCfg<NT>::prec, double2> GC;
    GB.e = GB.e1 = GC.e = GC.e1 = 0;
    const int dd = R.u - D * (R.u / D);
    SysScalars sc;
    if (SG && threadIdx.x >= kTpThreads) {
        // staging wavefronts: the copy, issued last before the first barrier
        // so that no wait of the prologue's falls between the two
        tp_row_constants<D, TB>(a, T, R);
        tp_seg_constants<D, TB>(a, T, R);
        sc = state_load(k.st_in + T.sys);
        if constexpr (SG && PC != CYMG_PRECOND_NONE) tp_stage_blocks<TB>(k, PcBlocks<PC>::get(k), T, stage);
        __syncthreads();
    } else {
        if (R.own) {
            vr = k.r[g]; vv = k.v[g]; vx = k.x[g]; vrt = k.rt[g];
            q2 = k.g2[g];
            vph = k.g1b[g].z;
        }
        tp_extents<D>(a, T, R.u, R.act, R.split, R.h, g, GB, GC);
        gath_entries(GB, a.Bcol, a.Bval);
        gath_entries(GC, a.Ccol, a.Cval);
        gath_values(GB, rec, (uint32_t)D, (uint32_t)dd);
        gath_values(GC, rec, 1u, 0u);
        tp_row_constants<D, TB>(a, T, R);
        tp_seg_constants<D, TB>(a, T, R);
        sc = state_load(k.st_in + T.sys);
        tile_tables<PC != CYMG_PRECOND_NONE>(a, L, T, k.blkoff, k.useblk);
        if (T.ws && threadIdx.x < T.nft) L.fsl[threadIdx.x] = k.ta.fslot[T.fb + threadIdx.x];
        tq_partials(k, T, L.bc);
        __syncthreads();
    }
    const uint32_t nt = k.ta.ntiles;
    const bool first = T.t == T.tb;
    if (!sc.done) tq_scalars<PC>(sc, L.bc);
    if (first && threadIdx.x == 0) state_store(k.st_out + T.sys, sc);
    if (sc.done) {
        tp_stage_drain<SG>();
        return;
    }
    tp_after_tables<D, TB, TpCfg<NT>::blk, SG>(a, L, T, PcBlocks<PC>::get(k), R);
    if constexpr (TpCfg<NT>::work == kTpThreads) tp_pcr_prefetch<D, TB>(a, L, T, R);
    double smax = 0.0, si = 0.0, shi = 0.0;
    if (R.own) {
        si = vr + (-sc.alpha) * vv;
        k.r[g] = si;
        k.x[g] = vx + sc.alpha * vph;
        smax = fabs(si);
        shi = q2.x + (-sc.alpha) * q2.y;
        L.x[R.u] = shi;
    }
    const ShatComb X{ sc.alpha };
    const double bsum = gath_sum(GB, rec, a.Bcol, a.Bval, (uint32_t)D, (uint32_t)dd, X);
    const double csum = gath_sum(GC, rec, a.Ccol, a.Cval, 1u, 0u, X);
    if (R.split && R.act && R.h == 1) {
        L.hb[R.u] = bsum;
        L.hc[R.u] = csum;
    }
    tp_block_prefetch<D, TB>(L, PcBlocks<PC>::get(k), R);
    __syncthreads();
    tp_operator<D, TB, TpCfg<NT>::blk, TpCfg<NT>::work == kTpThreads, SG>(L, T, R, bsum, csum);
    double tt = 0, ts = 0, rts = 0, rtt = 0;
    if (R.own) {
        const double ti = L.y[R.u];
        k.t[g] = ti;
        tt = ti * ti;
        ts = ti * si;
        rts = vrt * si;
        rtt = vrt * ti;
    }
    smax = wave_max(smax);
    tt = wave_sum(tt);
    ts = wave_sum(ts);
    rts = wave_sum(rts);
    rtt = wave_sum(rtt);
    if ((threadIdx.x & (kWave - 1)) == 0 && threadIdx.x < TpCfg<NT>::work) {
        const int wv = threadIdx.x / kWave;
        L.wr[wv] = smax;
        L.wr[NW + wv] = tt;
        L.wr[2 * NW + wv] = ts;
        L.wr[3 * NW + wv] = rts;
        L.wr[4 * NW + wv] = rtt;
    }
    const double wi = tp_block_apply<D, TB, TpCfg<NT>::blk, TpCfg<NT>::chunk, SG>(
        L, T, R, PcBlocks<PC>::get(k), L.y, reinterpret_cast<const TB*>(stage));
    if (R.own) reinterpret_cast<double2*>(k.g1b + g)[0] = make_double2(shi, wi);
    if (threadIdx.x == 0) {
        double m = L.wr[0], s0 = L.wr[NW], s1 = L.wr[2 * NW], s2 = L.wr[3 * NW], s3 = L.wr[4 * NW];
#pragma unroll
        for (int q = 1; q < NW; ++q) {
            m = nan_max(m, L.wr[q]);
            s0 += L.wr[NW + q];
            s1 += L.wr[2 * NW + q];
            s2 += L.wr[3 * NW + q];
            s3 += L.wr[4 * NW + q];
        }
        k.part[P_SMAX * nt + T.t] = m;
        *reinterpret_cast<double4*>(k.part4 + 4 * (size_t)T.t) = make_double4(s0, s1, s2, s3);
    }
}

static KArgs make_kargs(Context& c)
{
    KArgs k;
    std::memset(&k, 0, sizeof(k));
    k.fa = make_fiber_args(c);
    k.f32apply = c.fp32_apply;
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
    k.g1a = k.g1b = nullptr;
    k.g2 = nullptr;
    k.part = c.d_part.get();
    k.part4 = c.d_part.get() + (size_t)P_TT * c.nmec();
    k.fpart = c.d_part.get();
    k.ta.desc = c.d_tile_desc.get();
    k.ta.fslot = c.d_fslot.get();
    k.ta.ntiles = c.ntiles;
    k.st_in = nullptr;
    k.st_out = nullptr;
    k.nfib = c.nmec();
    k.nsys = c.nsys;
    k.nfiber = c.nfib;
    k.sa = make_solid_args(c);
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
    const dim3 g(c.nmec()), gf(c.nfib), b(kWave), b4(kApplyThreads), gs(c.nsys), gso(c.nsol), bso(kSolidOpThreads);
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
            if (c.nfib) hipLaunchKernelGGL((k_iter_b<D, PC>), gf, b, 0, c.stream, kB);
            if (c.nsol) hipLaunchKernelGGL((k_solid_phase<D, PC, ST_B>), gso, bso, 0, c.stream, kB);
            hipLaunchKernelGGL((k_iter_c<D, PC>), g, b4, 0, c.stream, kC);
            if (c.nfib) hipLaunchKernelGGL((k_iter_d<D, PC>), gf, b, 0, c.stream, kD);
            if (c.nsol) hipLaunchKernelGGL((k_solid_phase<D, PC, ST_D>), gso, bso, 0, c.stream, kD);
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
    // the next solve's first chunk follows a converged solve only: an attempt
    // that ran to its iteration limit would queue that many launches ahead
    predicted = all_conv ? c.h_scal[worst].iter : 0;
}

// Captures (or reuses) the graphs of 2 and 8 iterations of the two-launch
// iteration for these launch arguments.  Only kernel launches are captured.
template <int D, int PC, int NT>
static void iteration_graphs(Context& c, const KArgs& P0, const KArgs& Q0, const KArgs& P1, const KArgs& Q1,
                             dim3 g, dim3 bt)
{
    std::vector<unsigned char> key(4 * sizeof(KArgs) + 2 * sizeof(dim3));
    unsigned char* w = key.data();
    for (const KArgs* a : { &P0, &Q0, &P1, &Q1 }) {
        std::memcpy(w, a, sizeof(KArgs));
        w += sizeof(KArgs);
    }
    std::memcpy(w, &g, sizeof(dim3));
    std::memcpy(w + sizeof(dim3), &bt, sizeof(dim3));
    const int iters[2] = { 2, 8 };
    for (int v = 0; v < 2; ++v) {
        hipGraphExec_t& exec = c.iter_graph[PC][v];
        if (exec && c.iter_graph_key[PC][v] == key) continue;
        if (exec) {
            CYMG_HIP(hipGraphExecDestroy(exec));
            exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        CYMG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        for (int it = 0; it < iters[v]; it += 2) {
            hipLaunchKernelGGL((k_tp<D, PC, false, NT>), g, bt, 0, c.stream, P0);
            hipLaunchKernelGGL((k_tq<D, PC, NT>), g, bt, 0, c.stream, Q0);
            hipLaunchKernelGGL((k_tp<D, PC, false, NT>), g, bt, 0, c.stream, P1);
            hipLaunchKernelGGL((k_tq<D, PC, NT>), g, bt, 0, c.stream, Q1);
        }
        const hipError_t launch_err = hipGetLastError();
        const hipError_t end_err = hipStreamEndCapture(c.stream, &graph);
        CYMG_HIP(launch_err);
        CYMG_HIP(end_err);
        const hipError_t inst_err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        CYMG_HIP(inst_err);
        c.iter_graph_key[PC][v] = key;
    }
}

template <int D, int PC, int NT>
static void run_tiled(Context& c, const double* guess, int max_iter, double tol, int tol_mode,
                      cymg_solve_report* rep)
{
    KArgs k = make_kargs(c);
    k.part = c.d_tpart.get();
    k.part4 = c.d_tpart.get() + (size_t)P_TT * c.ntiles;
    SysScalars* S0 = c.d_scal.get();
    SysScalars* S1 = S0 + c.nsys;
    double4* g1[2] = { c.k_g1[0].get(), c.k_g1[1].get() };
    auto with = [&](SysScalars* in, SysScalars* out, int par) {
        KArgs q = k;
        q.st_in = in;
        q.st_out = out;
        q.g1a = g1[par];
        q.g1b = g1[par ^ 1];
        q.g2 = c.k_g2.get();
        return q;
    };
    // P reads phat/z of its parity and writes the other; Q reads what P wrote
    const KArgs kI = with(S0, S1, 0);
    const KArgs kP0 = with(S1, S0, 0), kQ0 = with(S0, S1, 0);
    const KArgs kP1 = with(S1, S0, 1), kQ1 = with(S0, S1, 1);
    const dim3 g(c.ntiles), b(kTileThreads), bt(NT);
    hipLaunchKernelGGL((k_tinit<D, PC>), g, b, 0, c.stream, kI, guess, tol, tol_mode, max_iter, c.prm.time_step);
    bool first_p = true;
    CYMG_HIP(hipGetLastError());
    // Chunks of iterations are queued one ahead: after queuing chunk k+1 the
    // host waits for the state snapshot that ends chunk k, so the device never
    // idles on the host's convergence check.  The first chunk covers 3/4 of the
    // previous solve's iterations; launches queued after convergence exit after
    // their prologue.
    int launched = 0;
    int& predicted = c.last_iters[PC];
    constexpr int kStep = 8;
    int first_chunk = predicted > 0 ? std::max(kStep, (predicted * 3) / 4) : kStep;
    first_chunk += first_chunk & 1;
    const size_t sbytes = sizeof(SysScalars) * c.nsys;
    SysScalars* snap[2] = { c.h_snap, c.h_snap + 4096 };
    // every iteration after the first is the same four launches with the same
    // arguments (the state lives in device memory), so chunks of 2 and 8
    // iterations are captured once as HIP graphs and replayed
    hipGraphExec_t* graphs = nullptr;
    if (c.use_graphs) {
        graphs = c.iter_graph[PC];
        iteration_graphs<D, PC, NT>(c, kP0, kQ0, kP1, kQ1, g, bt);
    }
    auto enqueue = [&](int n, int slot) {
        int it = 0;
        if (first_p) {
            hipLaunchKernelGGL((k_tp<D, PC, true, NT>), g, bt, 0, c.stream, kP0);
            hipLaunchKernelGGL((k_tq<D, PC, NT>), g, bt, 0, c.stream, kQ0);
            hipLaunchKernelGGL((k_tp<D, PC, false, NT>), g, bt, 0, c.stream, kP1);
            hipLaunchKernelGGL((k_tq<D, PC, NT>), g, bt, 0, c.stream, kQ1);
            first_p = false;
            it = 2;
        }
        if (graphs) {
            for (; it + 8 <= n; it += 8) CYMG_HIP(hipGraphLaunch(graphs[1], c.stream));
            for (; it + 2 <= n; it += 2) CYMG_HIP(hipGraphLaunch(graphs[0], c.stream));
        }
        for (; it < n; it += 2) {
            hipLaunchKernelGGL((k_tp<D, PC, false, NT>), g, bt, 0, c.stream, kP0);
            hipLaunchKernelGGL((k_tq<D, PC, NT>), g, bt, 0, c.stream, kQ0);
            hipLaunchKernelGGL((k_tp<D, PC, false, NT>), g, bt, 0, c.stream, kP1);
            hipLaunchKernelGGL((k_tq<D, PC, NT>), g, bt, 0, c.stream, kQ1);
        }
        CYMG_HIP(hipGetLastError());
        CYMG_HIP(hipMemcpyAsync(snap[slot], S1, sbytes, hipMemcpyDeviceToHost, c.stream));
        CYMG_HIP(hipEventRecord(c.chunk_ev[slot], c.stream));
        launched += n;
    };
    enqueue(first_chunk, 0);
    int cur = 0;
    while (true) {
        const bool more = launched <= max_iter + 4;
        if (more) enqueue(kStep, cur ^ 1);
        CYMG_HIP(hipEventSynchronize(c.chunk_ev[cur]));
        bool all = true;
        for (uint32_t s = 0; s < c.nsys; ++s) all = all && snap[cur][s].done;
        if (all || !more) break;
        cur ^= 1;
    }
    std::memcpy(c.h_scal, snap[cur], sbytes);
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
    // the next solve's first chunk follows a converged solve only: an attempt
    // that ran to its iteration limit would queue that many launches ahead
    predicted = all_conv ? c.h_scal[worst].iter : 0;
}

// Average duration of one launch of the two-launch iteration (P and Q
// alternately), from events around `pairs` iterations on the context's
// stream.  The iteration runs with a zero threshold so that every launch does
// its full work; an iteration that stops early (breakdown) is an error.
template <int D, int PC, int NT>
static double time_tiled(Context& c, int pairs)
{
    KArgs k = make_kargs(c);
    k.part = c.d_tpart.get();
    k.part4 = c.d_tpart.get() + (size_t)P_TT * c.ntiles;
    SysScalars* S0 = c.d_scal.get();
    SysScalars* S1 = S0 + c.nsys;
    double4* g1[2] = { c.k_g1[0].get(), c.k_g1[1].get() };
    auto with = [&](SysScalars* in, SysScalars* out, int par) {
        KArgs q = k;
        q.st_in = in;
        q.st_out = out;
        q.g1a = g1[par];
        q.g1b = g1[par ^ 1];
        q.g2 = c.k_g2.get();
        return q;
    };
    const KArgs kI = with(S0, S1, 0);
    const KArgs kP0 = with(S1, S0, 0), kQ0 = with(S0, S1, 0);
    const KArgs kP1 = with(S1, S0, 1), kQ1 = with(S0, S1, 1);
    const dim3 g(c.ntiles), b(kTileThreads), bt(NT);
    hipLaunchKernelGGL((k_tinit<D, PC>), g, b, 0, c.stream, kI, nullptr, 0.0, CYMG_TOL_RELATIVE, 1 << 30,
                       c.prm.time_step);
    pairs += pairs & 1;
    CYMG_HIP(hipEventRecord(c.ev[8], c.stream));
    for (int it = 0; it < pairs; it += 2) {
        if (it == 0) hipLaunchKernelGGL((k_tp<D, PC, true, NT>), g, bt, 0, c.stream, kP0);
        else hipLaunchKernelGGL((k_tp<D, PC, false, NT>), g, bt, 0, c.stream, kP0);
        hipLaunchKernelGGL((k_tq<D, PC, NT>), g, bt, 0, c.stream, kQ0);
        hipLaunchKernelGGL((k_tp<D, PC, false, NT>), g, bt, 0, c.stream, kP1);
        hipLaunchKernelGGL((k_tq<D, PC, NT>), g, bt, 0, c.stream, kQ1);
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
    const bool fresh = c.k_g2.cap < n;
    c.d_tpart.resize((size_t)P_SLOTS * (c.ntiles ? c.ntiles : 1));
    c.k_r.resize(n); c.k_rt.resize(n); c.k_v.resize(n); c.k_t.resize(n);
    c.k_g1[0].resize(n); c.k_g1[1].resize(n); c.k_g2.resize(n);
    if (fresh) {
        // the first launch reads phat_old, z_old and t only where beta = 0 or omega
        // multiplies them by zero: keep them finite
        CYMG_HIP(hipMemsetAsync(c.k_g1[0].get(), 0, c.k_g1[0].cap * sizeof(double4), c.stream));
        CYMG_HIP(hipMemsetAsync(c.k_g1[1].get(), 0, c.k_g1[1].cap * sizeof(double4), c.stream));
        CYMG_HIP(hipMemsetAsync(c.k_g2.get(), 0, c.k_g2.cap * sizeof(double2), c.stream));
        CYMG_HIP(hipMemsetAsync(c.k_t.get(), 0, c.k_t.cap * sizeof(double), c.stream));
    }
}

// the workgroup of the two-launch iteration: 256 threads when asked for and
// every tile fits (no Solid tile, every fiber within 256/DIM points), else 512
int tile_threads_of(const Context& c)
{
    if (c.tile_threads == 256 && (c.nsol == 0 || c.solid_tiles) &&
        c.largest <= (uint32_t)TpCfg<256>::max_points(c.dim))
        return 256;
    return 512;
}

// staged launches (kTpStaged threads) for the 512-thread tiles of a
// fiber-only layout: with FP32 blocks when CYMG_OPT_BLOCK_STAGE is 1 (the
// default), with FP32 or FP64 blocks when it is 2
static bool staged(const Context& c, int precond)
{
    const bool want = c.block_stage == 2 || (c.block_stage == 1 && precond == CYMG_PRECOND_FP32);
    return want && precond != CYMG_PRECOND_NONE && c.nsol == 0 && tile_threads_of(c) == kTpThreads;
}

// the two-launch iteration takes fibers of at most TpCfg<512>::max_points
// points (a tile of one fiber, a thread per unknown); a system with a longer
// fiber runs the four-launch iteration
bool tiles_fit(const Context& c)
{
    return c.largest <= (uint32_t)TpCfg<512>::max_points(c.dim);
}

// the iteration krylov_solve runs: two launches per iteration over tiles, or four
bool two_launch(const Context& c)
{
    return c.krylov_fused && (!c.nsol || c.solid_tiles) && tiles_fit(c);
}

int tile_max_points(const Context& c, int threads)
{
    return threads == 256 ? TpCfg<256>::max_points(c.dim) : TpCfg<512>::max_points(c.dim);
}

double krylov_time_launch(Context& c, int precond, int launches)
{
    if (c.nsol && !c.solid_tiles)
        throw Error(CYMG_ERR_UNSUPPORTED, "a Solid is too large for the two-launch iteration");
    if (!tiles_fit(c)) throw Error(CYMG_ERR_UNSUPPORTED, "a fiber is too long for the two-launch iteration");
    tiled_buffers(c);
    const int pairs = std::max(1, launches / 2);
    const bool small = tile_threads_of(c) == 256;
    const bool sg = staged(c, precond);
    auto go = [&](auto dim) -> double {
        constexpr int D = decltype(dim)::value;
        if (small) {
            if (precond == CYMG_PRECOND_FP64) return time_tiled<D, CYMG_PRECOND_FP64, 256>(c, pairs);
            if (precond == CYMG_PRECOND_FP32) return time_tiled<D, CYMG_PRECOND_FP32, 256>(c, pairs);
            return time_tiled<D, CYMG_PRECOND_NONE, 256>(c, pairs);
        }
        if (sg) {
            if (precond == CYMG_PRECOND_FP64) return time_tiled<D, CYMG_PRECOND_FP64, kTpStaged>(c, pairs);
            return time_tiled<D, CYMG_PRECOND_FP32, kTpStaged>(c, pairs);
        }
        if (precond == CYMG_PRECOND_FP64) return time_tiled<D, CYMG_PRECOND_FP64, 512>(c, pairs);
        if (precond == CYMG_PRECOND_FP32) return time_tiled<D, CYMG_PRECOND_FP32, 512>(c, pairs);
        return time_tiled<D, CYMG_PRECOND_NONE, 512>(c, pairs);
    };
    if (c.dim == 3) return go(std::integral_constant<int, 3>());
    return go(std::integral_constant<int, 2>());
}

void krylov_solve(Context& c, int precond, const double* guess_dev, int max_iter, double tol, int tol_mode,
                  cymg_solve_report* rep)
{
    // tiles hold whole fibers or one Solid of at most kSolidTileRows unknowns;
    // a larger Solid or a longer fiber runs the four-launch iteration, and so
    // do FP32 blocks applied in FP32 arithmetic (a nonlinear M^-1, which the
    // two-launch recurrences cannot carry)
    if (two_launch(c) && !(precond == CYMG_PRECOND_FP32 && c.fp32_apply)) {
        tiled_buffers(c);
        const bool small = tile_threads_of(c) == 256;
        const bool sg = staged(c, precond);
        auto go = [&](auto dim, auto nt) {
            constexpr int D = decltype(dim)::value, NT = decltype(nt)::value;
            if (precond == CYMG_PRECOND_FP64) run_tiled<D, CYMG_PRECOND_FP64, NT>(c, guess_dev, max_iter, tol, tol_mode, rep);
            else if (precond == CYMG_PRECOND_FP32) run_tiled<D, CYMG_PRECOND_FP32, NT>(c, guess_dev, max_iter, tol, tol_mode, rep);
            else if constexpr (NT != kTpStaged) run_tiled<D, CYMG_PRECOND_NONE, NT>(c, guess_dev, max_iter, tol, tol_mode, rep);
        };
        if (c.dim == 3) {
            if (small) go(std::integral_constant<int, 3>(), std::integral_constant<int, 256>());
            else if (sg) go(std::integral_constant<int, 3>(), std::integral_constant<int, kTpStaged>());
            else go(std::integral_constant<int, 3>(), std::integral_constant<int, 512>());
        } else {
            if (small) go(std::integral_constant<int, 2>(), std::integral_constant<int, 256>());
            else if (sg) go(std::integral_constant<int, 2>(), std::integral_constant<int, kTpStaged>());
            else go(std::integral_constant<int, 2>(), std::integral_constant<int, 512>());
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
