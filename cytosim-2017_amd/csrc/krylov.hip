// Preconditioned BiCGStab of the implicit step (Solver::BCGSP,
// bicgstab.h:247-338; Solver::BCGS, bicgstab.h:163-243; Monitor,
// bicgstab.h:51-92 and bicgstab.cc:9-49), entirely on the device.
//
// One iteration is four fused per-fiber launches:
//   A: [r = s - w t; x += w shat] (previous iteration), |r|_inf, convergence
//      test, p = r + beta (p - w v), phat = Binv p
//   B: v = M phat, <rt, v>  -> alpha = rho / <rt, v>
//   C: s = r - alpha v, x += alpha phat, |s|_inf, shat = Binv s
//   D: t = M shat, <t,t>, <t,s>, <rt,s>, <rt,t>  -> omega, next rho, beta
// The next rho is <rt, r_new> = <rt, s> - omega <rt, t>, which needs no extra
// pass over r.  Every reduction is per system: partials per fiber, summed in
// fiber order by the last workgroup of the system, so results do not depend on
// workgroup scheduling.  Scalars and the stop flag stay in HBM; the host only
// reads the flag between chunks of iterations.
#include "args.h"
#include "pc_ops.h"

namespace cymg {

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
    double* part;
    unsigned* counter;
    SysScalars* scal;
    uint32_t nfib;
};

enum { P_RR = 0, P_RMAX = 1, P_DELTA = 2, P_SMAX = 3, P_TT = 4, P_TS = 5, P_RTS = 6, P_RTT = 7, P_SLOTS = 8 };

template <int PC>
__device__ __forceinline__ void apply_block(const KArgs& k, uint32_t f, int bs, const double* xs, double* y)
{
    if constexpr (PC == CYMG_PRECOND_FP64)
        block_apply<double>(k.B64 + k.blkoff[f], bs, k.useblk[f] != 0, xs, xs, y);
    else if constexpr (PC == CYMG_PRECOND_FP32)
        block_apply<float>(k.B32 + k.blkoff[f], bs, k.useblk[f] != 0, xs, xs, y);
}

// r = b - M x0 (or b), rt = r, x = x0 (or 0); rho = <rt, r>
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
    const uint32_t sys = a.fib_sys[f];
    const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
    if (last_of_system(k.counter, sys, fe - fb)) {
        const double dot = system_sum(k.part + P_RR * k.nfib, fb, fe);
        if (threadIdx.x == 0) {
            SysScalars& sc = k.scal[sys];
            sc.rho2 = 1.0;
            sc.rho1 = dot;
            sc.alpha = 0.0;
            sc.omega = 1.0;
            sc.beta = (sc.rho1 / sc.rho2) * (sc.alpha / sc.omega);
            sc.tol = (tol_mode == CYMG_TOL_RELATIVE) ? tolerance * sc.bnorm : tolerance * (sc.noise * dt);
            sc.resid = __builtin_inf();
            sc.resid_s = __builtin_inf();
            sc.iter = 0;
            sc.iter_max = iter_max;
            sc.done = 0;
            sc.flag = 0;
            sc.pending = 0;
        }
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kWave) void k_iter_a(KArgs k)
{
    __shared__ FiberLds<D> s;
    const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const uint32_t sys = a.fib_sys[f];
    const SysScalars sc = k.scal[sys];
    if (sc.done) return;
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int n = D * (int)np;
    double rmax = 0.0;
    for (int i = threadIdx.x; i < n; i += kWave) {
        double ri = k.r[o + i];
        if (sc.pending) {
            // x += omega shat (BCGS: shat = s), r = s - omega t
            const double sh = PC ? k.sh[o + i] : ri;
            k.x[o + i] = k.x[o + i] + sc.omega * sh;
            ri = ri + (-sc.omega) * k.t[o + i];
            k.r[o + i] = ri;
        }
        rmax = nan_max(rmax, fabs(ri));
        double pi;
        if (sc.beta == 0.0) pi = ri;
        else pi = sc.beta * (k.p[o + i] + (-sc.omega) * k.v[o + i]) + ri;
        k.p[o + i] = pi;
        s.x[i] = pi;
    }
    if constexpr (PC != CYMG_PRECOND_NONE) {
        __syncthreads();
        apply_block<PC>(k, f, n, s.x, k.ph + o);
    }
    rmax = wave_max(rmax);
    if (threadIdx.x == 0) k.part[P_RMAX * k.nfib + f] = rmax;
    const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
    if (last_of_system(k.counter, sys, fe - fb)) {
        double res = system_max(k.part + P_RMAX * k.nfib, fb, fe);
        if (threadIdx.x == 0) {
            SysScalars& w = k.scal[sys];
            w.pending = 0;
            if (res != res) {
                w.resid = __builtin_inf();
                w.done = 1;
            } else {
                w.resid = res;
                if (w.iter > w.iter_max) w.done = 1;
                else if (res < w.tol) w.done = 1;
                else if (w.rho1 == 0.0) { w.done = 1; w.flag = 2; }
            }
        }
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kWave) void k_iter_b(KArgs k)
{
    __shared__ FiberLds<D> s;
    const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const uint32_t sys = a.fib_sys[f];
    if (k.scal[sys].done) return;
    const double* X = PC ? k.ph : k.p;
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int n = D * (int)np;
    for (int i = threadIdx.x; i < n; i += kWave) s.x[i] = X[o + i];
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
    const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
    if (last_of_system(k.counter, sys, fe - fb)) {
        const double delta = system_sum(k.part + P_DELTA * k.nfib, fb, fe);
        if (threadIdx.x == 0) {
            SysScalars& w = k.scal[sys];
            if (PC != CYMG_PRECOND_NONE && delta == 0.0) { w.done = 1; w.flag = 4; }
            else w.alpha = w.rho1 / delta;
        }
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kWave) void k_iter_c(KArgs k)
{
    __shared__ FiberLds<D> s;
    const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const uint32_t sys = a.fib_sys[f];
    const SysScalars sc = k.scal[sys];
    if (sc.done) return;
    const double* X = PC ? k.ph : k.p;
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int n = D * (int)np;
    double smax = 0.0;
    for (int i = threadIdx.x; i < n; i += kWave) {
        const double si = k.r[o + i] + (-sc.alpha) * k.v[o + i];
        k.r[o + i] = si;
        k.x[o + i] = k.x[o + i] + sc.alpha * X[o + i];
        smax = nan_max(smax, fabs(si));
        s.x[i] = si;
    }
    if constexpr (PC != CYMG_PRECOND_NONE) {
        __syncthreads();
        apply_block<PC>(k, f, n, s.x, k.sh + o);
    }
    smax = wave_max(smax);
    if (threadIdx.x == 0) k.part[P_SMAX * k.nfib + f] = smax;
    const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
    if (last_of_system(k.counter, sys, fe - fb)) {
        const double m = system_max(k.part + P_SMAX * k.nfib, fb, fe);
        if (threadIdx.x == 0) k.scal[sys].resid_s = (m != m) ? __builtin_inf() : m;
    }
}

template <int D, int PC>
__global__ __launch_bounds__(kWave) void k_iter_d(KArgs k)
{
    __shared__ FiberLds<D> s;
    const uint32_t f = blockIdx.x;
    const FiberArgs& a = k.fa;
    const uint32_t sys = a.fib_sys[f];
    if (k.scal[sys].done) return;
    const double* X = PC ? k.sh : k.r;
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int n = D * (int)np;
    for (int i = threadIdx.x; i < n; i += kWave) s.x[i] = X[o + i];
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
    if (threadIdx.x == 0) {
        k.part[P_TT * k.nfib + f] = tt;
        k.part[P_TS * k.nfib + f] = ts;
        k.part[P_RTS * k.nfib + f] = rts;
        k.part[P_RTT * k.nfib + f] = rtt;
    }
    const uint32_t fb = a.sys_begin[sys], fe = a.sys_begin[sys + 1];
    if (last_of_system(k.counter, sys, fe - fb)) {
        const double Stt = system_sum(k.part + P_TT * k.nfib, fb, fe);
        const double Sts = system_sum(k.part + P_TS * k.nfib, fb, fe);
        const double Srts = system_sum(k.part + P_RTS * k.nfib, fb, fe);
        const double Srtt = system_sum(k.part + P_RTT * k.nfib, fb, fe);
        if (threadIdx.x == 0) {
            SysScalars& w = k.scal[sys];
            if (Stt == 0.0) {
                w.done = 1; w.flag = 0; w.resid = w.resid_s;
            } else {
                const double om = Sts / Stt;
                if (om == 0.0) {
                    w.done = 1; w.flag = 3; w.resid = w.resid_s;
                } else {
                    w.omega = om;
                    w.pending = 1;
                    w.rho2 = w.rho1;
                    w.rho1 = Srts - om * Srtt;
                    w.beta = (w.rho1 / w.rho2) * (w.alpha / w.omega);
                    w.iter += 1;
                }
            }
        }
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
    k.part = c.d_part.get();
    k.counter = c.d_counter.get();
    k.scal = c.d_scal.get();
    k.nfib = c.nfib;
    return k;
}

template <int D, int PC>
static void run(Context& c, const double* guess, int max_iter, double tol, int tol_mode, cymg_solve_report* rep)
{
    KArgs k = make_kargs(c);
    const dim3 g(c.nfib), b(kWave);
    hipLaunchKernelGGL((k_init<D, PC>), g, b, 0, c.stream, k, guess, tol, tol_mode, max_iter, c.prm.time_step);
    CYMG_HIP(hipGetLastError());
    int launched = 0;
    int chunk = 4;
    const size_t sbytes = sizeof(SysScalars) * c.nsys;
    while (true) {
        for (int it = 0; it < chunk; ++it) {
            hipLaunchKernelGGL((k_iter_a<D, PC>), g, b, 0, c.stream, k);
            hipLaunchKernelGGL((k_iter_b<D, PC>), g, b, 0, c.stream, k);
            hipLaunchKernelGGL((k_iter_c<D, PC>), g, b, 0, c.stream, k);
            hipLaunchKernelGGL((k_iter_d<D, PC>), g, b, 0, c.stream, k);
        }
        CYMG_HIP(hipGetLastError());
        launched += chunk;
        // one more A pass decides convergence of the last iterate
        CYMG_HIP(hipMemcpyAsync(c.h_scal, c.d_scal.get(), sbytes, hipMemcpyDeviceToHost, c.stream));
        CYMG_HIP(hipStreamSynchronize(c.stream));
        bool all = true;
        for (uint32_t s = 0; s < c.nsys; ++s) all = all && c.h_scal[s].done;
        if (all) break;
        if (launched > max_iter + 2) {
            // iteration cap reached: a final A pass sets the flags
            hipLaunchKernelGGL((k_iter_a<D, PC>), g, b, 0, c.stream, k);
            CYMG_HIP(hipMemcpyAsync(c.h_scal, c.d_scal.get(), sbytes, hipMemcpyDeviceToHost, c.stream));
            CYMG_HIP(hipStreamSynchronize(c.stream));
            break;
        }
        if (chunk < 16) chunk *= 2;
    }
    // report system 0 (single-system contexts); batched callers read per system
    const SysScalars& w = c.h_scal[0];
    rep->iterations = w.iter;
    rep->flag = w.flag;
    rep->residual = w.resid;
    rep->threshold = w.tol;
    rep->converged = w.resid < w.tol;
    rep->noise_level = w.noise * c.prm.time_step;
    rep->precond_used = PC;
}

void krylov_solve(Context& c, int precond, const double* guess_dev, int max_iter, double tol, int tol_mode,
                  cymg_solve_report* rep)
{
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
