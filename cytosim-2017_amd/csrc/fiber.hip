// Per-fiber kernels of the implicit step: Meca::prepare, the device-side
// emitters, the force/RHS setup, the standalone operator, and the post-solve
// update (vPTS += vSOL, forces, getPoints reshape).  One wavefront per fiber.
#include <algorithm>

#include "args.h"

namespace cymg {

// launch kern<D, P> with the LDS capacity P the layout needs: the short one
// unless a fiber is longer (fewer resident workgroups otherwise)
#define CYMG_FIBER_LAUNCH(kern, D, ...)                                                    \
    do {                                                                                   \
        if (c.largest <= (uint32_t)kShortFiberPoints)                                      \
            hipLaunchKernelGGL((kern<D, kShortFiberPoints>), __VA_ARGS__);                 \
        else                                                                               \
            hipLaunchKernelGGL((kern<D, kMaxFiberPoints>), __VA_ARGS__);                   \
    } while (0)

constexpr double kRealEpsilon = 128.0 * 2.220446049250313e-16;  // src/math/real.h:32

// ---------------------------------------------------------------------------
// Meca::prepare (meca.cc:488-537): putPoints, then Fiber::prepareMecable
// (fiber.cc:450-460): setDragCoefficient (fiber.cc:315-363, 428-447),
// storeDifferences (rigid_fiber.cc:111-117), makeProjection
// (rigid_fiber_project.cc:50-87), rfRigidity = rigidity / fnCut^3.
// dragp: per fiber the hydrodynamic radii (0: the context's r0, r1) and the
// cylinder height of the surface drag model (0: drag in a volume).
template <int D, int P>
__global__ __launch_bounds__(kWave) void k_prepare(const uint32_t* __restrict__ npv, const uint32_t* __restrict__ idxv,
                                                   const uint32_t* __restrict__ segv, const double* __restrict__ cutv,
                                                   const double* __restrict__ rigv, const double* __restrict__ pos,
                                                   double* __restrict__ pts, double* __restrict__ mob,
                                                   double* __restrict__ rfrig, double* __restrict__ diff,
                                                   double* __restrict__ jd, double* __restrict__ je,
                                                   double* __restrict__ jd0, double* __restrict__ je0,
                                                   int* __restrict__ status, const double* __restrict__ dragp,
                                                   double visc, double r0d, double r1d, double* __restrict__ pcr,
                                                   uint32_t nseg)
{
    __shared__ double sdif[P * D];
    __shared__ double sd[P], se[P];
    const uint32_t f = blockIdx.x;
    const uint32_t np = npv[f], idx = idxv[f], seg = segv[f];
    const int S = (int)np - 1;
    const double cut = cutv[f];
    for (int i = threadIdx.x; i < (int)(D * np); i += kWave) pts[D * idx + i] = pos[D * idx + i];
    const double sc = 1.0 / cut;
    for (int i = threadIdx.x; i < D * S; i += kWave) {
        const double v = sc * (pos[D * idx + i + D] - pos[D * idx + i]);
        sdif[i] = v;
        diff[D * seg + i] = v;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < S; q += kWave) {
        double dd = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) dd += sdif[D * q + d] * sdif[D * q + d];
        sd[q] = 2 * dd;
        jd0[seg + q] = 2 * dd;
        if (q + 1 < S) {
            double od = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) od += sdif[D * q + d] * sdif[D * q + D + d];
            se[q] = -od;
            je0[seg + q] = -od;
        }
    }
    __syncthreads();
    if (S <= kWave) {
        // The reduction coefficients of parallel cyclic reduction on JJ' (the
        // recurrence of pcr_solve_span on the rows a, b, c; they do not depend
        // on the right-hand side), so that each projection during the solve
        // only combines right-hand sides.  Rows past the fiber are identity rows.
        const int q = threadIdx.x;
        double ra = 0.0, rb = 1.0, rc = 0.0;
        if (q < S) {
            rb = sd[q];
            if (q > 0) ra = se[q - 1];
            if (q + 1 < S) rc = se[q];
        }
        for (int l = 0; l < kPcrLevels; ++l) {
            const int k = 1 << l;
            double al = __shfl_up(ra, k, kWave), bl = __shfl_up(rb, k, kWave), cl = __shfl_up(rc, k, kWave);
            double au = __shfl_down(ra, k, kWave), bu = __shfl_down(rb, k, kWave), cu = __shfl_down(rc, k, kWave);
            if (q < k) { al = 0.0; bl = 1.0; cl = 0.0; }
            if (q + k >= kWave) { au = 0.0; bu = 1.0; cu = 0.0; }
            const double alpha = -ra / bl;
            const double gamma = -rc / bu;
            rb = rb + alpha * cl + gamma * au;
            ra = alpha * al;
            rc = gamma * cu;
            if (q < S) {
                pcr[(size_t)(2 * l) * nseg + seg + q] = alpha;
                pcr[(size_t)(2 * l + 1) * nseg + seg + q] = gamma;
            }
        }
        if (q < S) pcr[(size_t)(2 * kPcrLevels) * nseg + seg + q] = rb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // dpttrf
        int info = 0;
        for (int i = 0; i < S - 1; ++i) {
            if (sd[i] <= 0.0) { info = i + 1; break; }
            const double ei = se[i];
            se[i] = ei / sd[i];
            sd[i + 1] -= se[i] * ei;
        }
        if (!info && S > 0 && sd[S - 1] <= 0.0) info = S;
        if (info) atomicAdd(status, 1);
        // drag and mobility
        const double r0 = dragp[3 * f] > 0 ? dragp[3 * f] : r0d;
        const double r1 = dragp[3 * f + 1] > 0 ? dragp[3 * f + 1] : r1d;
        const double height = dragp[3 * f + 2];
        const double len = (double)S * cut;
        double drag;
        if (height > 0) {
            // Fiber::dragCoefficientSurface (fiber.cc:401-414)
            drag = 2 * M_PI * visc * len / acosh(1 + height / r0);
        } else {
            // Fiber::dragCoefficientVolume (fiber.cc:315-363)
            double lenc = len > r1 ? r1 : len;
            const double drag_sphere = 6 * r0;
            const double drag_cyl = 3.0 * len / (log(0.5 * lenc / r0) + 0.312);
            drag = M_PI * visc * (drag_cyl > drag_sphere ? drag_cyl : drag_sphere);
        }
        mob[f] = (double)np / drag;
        rfrig[f] = rigv[f] / (cut * cut * cut);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < S; q += kWave) {
        jd[seg + q] = sd[q];
        if (q + 1 < S) je[seg + q] = se[q];
    }
}

void launch_prepare(Context& c)
{
    c.d_status.resize(4);
    CYMG_HIP(hipMemsetAsync(c.d_status.get(), 0, 4 * sizeof(int), c.stream));
    launch_solid_prepare(c);
    if (c.nfib) launch_fiber_prepare(c);
    CYMG_HIP(hipMemcpyAsync(c.h_status, c.d_status.get(), 4 * sizeof(int), hipMemcpyDeviceToHost, c.stream));
}

void launch_fiber_prepare(Context& c)
{
    const dim3 g(c.nfib), b(kWave);
    if (c.dim == 3)
        CYMG_FIBER_LAUNCH(k_prepare, 3, g, b, 0, c.stream, c.d_np.get(), c.d_idx.get(), c.d_seg.get(), c.d_cut.get(),
                           c.d_rig.get(), c.d_pos.get(), c.d_pts.get(), c.d_mob.get(), c.d_rfrig.get(),
                           c.d_diff.get(), c.d_jd.get(), c.d_je.get(), c.d_jd0.get(), c.d_je0.get(), c.d_status.get(), c.d_drag.get(),
                           c.prm.viscosity, c.prm.hydro_radius0, c.prm.hydro_radius1, c.d_pcr.get(), c.nseg);
    else
        CYMG_FIBER_LAUNCH(k_prepare, 2, g, b, 0, c.stream, c.d_np.get(), c.d_idx.get(), c.d_seg.get(), c.d_cut.get(),
                           c.d_rig.get(), c.d_pos.get(), c.d_pts.get(), c.d_mob.get(), c.d_rfrig.get(),
                           c.d_diff.get(), c.d_jd.get(), c.d_je.get(), c.d_jd0.get(), c.d_je0.get(), c.d_status.get(), c.d_drag.get(),
                           c.prm.viscosity, c.prm.hydro_radius0, c.prm.hydro_radius1, c.d_pcr.get(), c.nseg);
    CYMG_HIP(hipGetLastError());
}

// dst[point] = src[map[point]] for every coordinate, 0 where map is invalid:
// the values a new registration inherits from the previous one
__global__ void k_gather_points(uint32_t n, int D, const uint32_t* __restrict__ map, const double* __restrict__ src,
                                double* __restrict__ dst)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = i / (uint32_t)D, m = map[p];
    dst[i] = m == kInvalid32 ? 0.0 : src[(size_t)D * m + (i - (uint32_t)D * p)];
}

void launch_gather_points(Context& c, const uint32_t* map, const double* src, double* dst, uint32_t npts)
{
    const uint32_t n = (uint32_t)c.dim * npts;
    if (!n) return;
    hipLaunchKernelGGL(k_gather_points, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, c.dim, map, src, dst);
    CYMG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Fiber::setInteractions, CONFINE_INSIDE (fiber.cc:471-481) with
// SpaceSphere::inside (space_sphere.cc:113-116) and SpaceSphere::setInteraction
// -> Meca::interLongClamp(pe, 0, radius, stiffness) (space_sphere.cc:148-151,
// meca_inter.cc:1977-2022).  Fixed slots per point: D(D+1)/2 mC entries and D
// vBAS entries, invalid keys for points inside.
__device__ __forceinline__ uint64_t lower_key(uint32_t i, uint32_t j)
{
    const uint32_t r = i > j ? i : j, c = i > j ? j : i;
    return ((uint64_t)c << 32) | r;
}

template <int D>
__global__ void k_emit_confine(uint32_t npts, const double* __restrict__ pos, double radius, double w,
                               uint64_t* __restrict__ keyC, double* __restrict__ valC,
                               uint64_t* __restrict__ keyBas, double* __restrict__ valBas)
{
#pragma clang fp contract(off)
    constexpr int NC = D * (D + 1) / 2;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    double x[D];
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        x[d] = pos[D * p + d];
        r2 += x[d] * x[d];
    }
    uint64_t* kc = keyC + (size_t)NC * p;
    double* vc = valC + (size_t)NC * p;
    uint64_t* kb = keyBas + (size_t)D * p;
    double* vb = valBas + (size_t)D * p;
    bool emit = !(r2 <= radius * radius);
    double axis[D], an = 0.0;
    if (emit) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            axis[d] = x[d] - 0.0;
            an += axis[d] * axis[d];
        }
        an = sqrt(an);
        if (an < kRealEpsilon) emit = false;
    }
    if (!emit) {
        for (int k = 0; k < NC; ++k) { kc[k] = kInvalidKey; vc[k] = 0.0; }
        for (int k = 0; k < D; ++k) { kb[k] = kInvalidKey; vb[k] = 0.0; }
        return;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) axis[d] /= an;
    double ac = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) ac += axis[d] * 0.0;
    const uint32_t inx = D * p;
    int k = 0;
    double len = radius;
    if (len < an) {
        len /= an;
        const double facX = w * len * (an + ac);
        const double facC = w * (1.0 - len);
        for (int i = 0; i < D; ++i) {
            kc[k] = lower_key(inx + i, inx + i);
            vc[k] = w * (len * (1.0 - axis[i] * axis[i]) - 1.0);
            ++k;
            for (int j = i + 1; j < D; ++j) {
                kc[k] = lower_key(inx + i, inx + j);
                vc[k] = -(w * len * axis[i] * axis[j]);
                ++k;
            }
            kb[i] = inx + i;
            vb[i] = facX * axis[i] + facC * 0.0;
        }
    } else {
        const double facX = w * (len + ac);
        for (int i = 0; i < D; ++i) {
            for (int j = i; j < D; ++j) {
                kc[k] = lower_key(inx + i, inx + j);
                vc[k] = -(w * axis[i] * axis[j]);
                ++k;
            }
            kb[i] = inx + i;
            vb[i] = facX * axis[i];
        }
    }
}

void launch_emit_confine(Context& c, uint64_t* keyC, double* valC, uint64_t* keyBas, double* valBas, double radius,
                         double stiffness)
{
    // Fiber::setInteractions: the fibers' points only (a Solid is not confined, solid.cc:112-114)
    if (!c.fib_pts) return;
    const dim3 b(256), g((c.fib_pts + 255) / 256);
    if (c.dim == 3)
        hipLaunchKernelGGL(k_emit_confine<3>, g, b, 0, c.stream, c.fib_pts, c.d_pos.get(), radius, stiffness, keyC,
                           valC, keyBas, valBas);
    else
        hipLaunchKernelGGL(k_emit_confine<2>, g, b, 0, c.stream, c.fib_pts, c.d_pos.get(), radius, stiffness, keyC,
                           valC, keyBas, valBas);
    CYMG_HIP(hipGetLastError());
}

// The emitters evaluate every product and sum rounded as written
// (contraction off), the arithmetic of the reference source.
// Couple::setInteractions (couple.cc:87-92) -> interLink(PI, PI)
// (meca_inter.cc:285-321) with FiberNaked::interpolate (fiber_naked.cc:1050-1065).
// 10 slots per couple; overlapping hands give invalid keys.
__device__ __forceinline__ void interp(uint32_t np, uint32_t idx, double cut, double ab, uint32_t& m1, double& co)
{
    if (ab <= 0) { m1 = idx; co = 0.0; return; }
    double n;
    const double fr = modf(ab / cut, &n);
    const uint32_t rd = (uint32_t)n;
    if (rd < np - 1) { m1 = idx + rd; co = fr; }
    else { m1 = idx + np - 2; co = 1.0; }
}

__global__ void k_emit_couples(uint32_t n, const uint32_t* __restrict__ fa, const double* __restrict__ aa,
                               const uint32_t* __restrict__ fb, const double* __restrict__ ab,
                               const double* __restrict__ wv, const uint8_t* __restrict__ state,
                               const uint32_t* __restrict__ npv, const uint32_t* __restrict__ idxv,
                               const double* __restrict__ cutv, uint64_t* __restrict__ keyB,
                               double* __restrict__ valB)
{
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t f1 = fa[k], f2 = fb[k];
    uint32_t a1, b1;
    double ca, cb;
    interp(npv[f1], idxv[f1], cutv[f1], aa[k], a1, ca);
    interp(npv[f2], idxv[f2], cutv[f2], ab[k], b1, cb);
    const uint32_t a2 = a1 + 1, b2 = b1 + 1;
    uint64_t* kk = keyB + (size_t)10 * k;
    double* vv = valB + (size_t)10 * k;
    // only bridging couples (both hands attached) interact (couple.cc:87-92)
    if ((state[k] & 3u) != 3u || (f1 == f2 && (a1 == b1 || a1 == b2 || a2 == b1 || a2 == b2))) {
        for (int i = 0; i < 10; ++i) { kk[i] = kInvalidKey; vv[i] = 0.0; }
        return;
    }
    const double w = wv[k];
    const double cc[4] = { 1.0 - ca, ca, -(1.0 - cb), -cb };
    const double cw[4] = { w * cc[0], w * cc[1], w * cc[2], w * cc[3] };
    const uint32_t inx[4] = { a1, a2, b1, b2 };
    int s = 0;
    for (int j = 0; j < 4; ++j)
        for (int i = j; i < 4; ++i) {
            kk[s] = lower_key(inx[i], inx[j]);
            vv[s] = -(cc[j] * cw[i]);
            ++s;
        }
}

// Aster::setInteractions (aster.cc:53-89): links between a Solid point and a
// fiber, 6 slots per link in the order of the host interLink calls:
//   kind 0: interLink(PointExact(solid, s1), fiber minus end)       (meca_inter.cc:196-230)
//   kind 1: interLink(fiber at abscissa ab, PointExact(solid, s1))  (meca_inter.cc:240-275)
//   kind 2: interLink(PointInterpolated(solid, s1, s2, ab), fiber plus end)
// The two ends of a link are on different Mecables, so no link is skipped.
__global__ void k_emit_asters(uint32_t n, const int* __restrict__ kind, const uint32_t* __restrict__ s1v,
                              const uint32_t* __restrict__ s2v, const uint32_t* __restrict__ fv,
                              const double* __restrict__ abv, const double* __restrict__ wv,
                              const uint32_t* __restrict__ npv, const uint32_t* __restrict__ idxv,
                              const double* __restrict__ cutv, uint64_t* __restrict__ keyB, double* __restrict__ valB)
{
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint64_t* kk = keyB + (size_t)6 * k;
    double* vv = valB + (size_t)6 * k;
    const double w = wv[k];
    const uint32_t f = fv[k];
    const int kd = kind[k];
    if (kd == 0) {
        const uint32_t a = s1v[k], b = idxv[f];
        kk[0] = lower_key(a, a); vv[0] = -w;
        kk[1] = lower_key(a, b); vv[1] = +w;
        kk[2] = lower_key(b, b); vv[2] = -w;
        for (int i = 3; i < 6; ++i) { kk[i] = kInvalidKey; vv[i] = 0.0; }
        return;
    }
    uint32_t a1, a2, b;
    double ca;
    if (kd == 1) {
        interp(npv[f], idxv[f], cutv[f], abv[k], a1, ca);
        a2 = a1 + 1;
        b = s1v[k];
    } else {
        a1 = s1v[k];
        a2 = s2v[k];
        ca = abv[k];
        b = idxv[f] + npv[f] - 1;
    }
    const double cc[3] = { 1.0 - ca, ca, -1.0 };
    const double cw[3] = { w * cc[0], w * cc[1], -w };
    const uint32_t inx[3] = { a1, a2, b };
    int s = 0;
    for (int q = 0; q < 3; ++q)
        for (int l = q; l < 3; ++l) {
            kk[s] = lower_key(inx[q], inx[l]);
            vv[s] = -(cc[q] * cw[l]);
            ++s;
        }
}

void launch_emit_asters(Context& c, uint64_t* keyB, double* valB)
{
    if (!c.nasters) return;
    const dim3 b(256), g((c.nasters + 255) / 256);
    hipLaunchKernelGGL(k_emit_asters, g, b, 0, c.stream, c.nasters, c.ast_kind.get(), c.ast_s1.get(), c.ast_s2.get(),
                       c.ast_f.get(), c.ast_ab.get(), c.ast_w.get(), c.d_np.get(), c.d_idx.get(), c.d_cut.get(),
                       keyB, valB);
    CYMG_HIP(hipGetLastError());
}

void launch_emit_couples(Context& c, uint64_t* keyB, double* valB)
{
    if (!c.ncouples) return;
    const dim3 b(256), g((c.ncouples + 255) / 256);
    hipLaunchKernelGGL(k_emit_couples, g, b, 0, c.stream, c.ncouples, c.cpl_fa.get(), c.cpl_aa.get(), c.cpl_fb.get(),
                       c.cpl_ab.get(), c.cpl_w.get(), c.cpl_state.get(), c.d_np.get(), c.d_idx.get(),
                       c.d_cut.get(), keyB, valB);
    CYMG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Couple::stepAA / stepAF / stepFA (couple.cc:125-150) for couples with motor
// hands, from the points the fibers hold after the solve (the reshaped points,
// FiberNaked::getPoints).  Couple::force1 (couple.cc:274-282)
// is f = k (pos2 - pos1); hand 1 steps under f and hand 2 under -f.
// Motor::stepLoaded (motor.cc:32-57): load = f . dirFiber, with dirFiber =
// (P[p+1] - P[p]) / fnCut at the hand's first interpolation point
// (fiber_naked.h:205), dabs = max_speed_dt + load * var_speed_dt, clamped to
// [min_dabs, max_dabs] when limit_speed (motor_prop.cc:51-65); an attached
// motor whose partner hand is free steps unloaded by max_speed_dt
// (motor.cc:19-28).  FiberBinder::moveBy / checkFiberRange
// (fiber_binder.cc:59-65, 111-118) and Hand::handleOutOfRange (hand.cc:102-122)
// for fibers that neither grow nor shrink: past an end the hand holds the end
// or detaches.  The unbinding tests (hand.cc:145-205) need a random stream and
// an unbinding rate; they are not modelled (unbinding off).
// state bits: 1 hand 1 attached, 2 hand 2 attached, 4 hand 1 is a motor,
// 8 hand 2 is a motor.
template <int D>
__device__ __forceinline__ void hand_geometry(const double* __restrict__ pts, uint32_t np, uint32_t idx, double cut,
                                              double ab, double* pos, double* dir)
{
#pragma clang fp contract(off)
    uint32_t m1;
    double co;
    interp(np, idx, cut, ab, m1, co);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double p1 = pts[D * m1 + d], p2 = pts[D * (m1 + 1) + d];
        pos[d] = p1 + co * (p2 - p1);
        dir[d] = (p2 - p1) / cut;
    }
}

__device__ __forceinline__ void move_hand(double& ab, uint32_t& st, uint32_t bit, double dabs, double len, int hold)
{
#pragma clang fp contract(off)
    ab = ab + dabs;
    if (ab < 0.0) {
        if (hold) ab = 0.0;
        else st &= ~bit;
    } else if (ab > len) {
        if (hold) ab = len;
        else st &= ~bit;
    }
}

template <int D>
__global__ void k_step_couples(uint32_t n, const uint32_t* __restrict__ fa, double* __restrict__ aa,
                               const uint32_t* __restrict__ fb, double* __restrict__ ab,
                               const double* __restrict__ wv, uint8_t* __restrict__ state,
                               const uint32_t* __restrict__ npv, const uint32_t* __restrict__ idxv,
                               const double* __restrict__ cutv, const double* __restrict__ pts, MotorStep ms)
{
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t st = state[k];
    if (!(st & 12u)) return;
    const uint32_t f1 = fa[k], f2 = fb[k];
    double a1 = aa[k], a2 = ab[k];
    const double len1 = cutv[f1] * (double)(npv[f1] - 1), len2 = cutv[f2] * (double)(npv[f2] - 1);
    if ((st & 3u) == 3u) {
        double p1[3], d1[3], p2[3], d2[3], f[3], g[3];
        hand_geometry<D>(pts, npv[f1], idxv[f1], cutv[f1], a1, p1, d1);
        hand_geometry<D>(pts, npv[f2], idxv[f2], cutv[f2], a2, p2, d2);
        const double w = wv[k];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            f[d] = w * (p2[d] - p1[d]);
            g[d] = -f[d];
        }
        if (st & 4u) {
            double load = f[0] * d1[0];
#pragma unroll
            for (int d = 1; d < D; ++d) load = load + f[d] * d1[d];
            double dabs = ms.max_speed_dt + load * ms.var_speed_dt;
            if (ms.limit_speed) dabs = dabs < ms.min_dabs ? ms.min_dabs : (dabs > ms.max_dabs ? ms.max_dabs : dabs);
            move_hand(a1, st, 1u, dabs, len1, ms.hold_end);
        }
        if (st & 8u) {
            double load = g[0] * d2[0];
#pragma unroll
            for (int d = 1; d < D; ++d) load = load + g[d] * d2[d];
            double dabs = ms.max_speed_dt + load * ms.var_speed_dt;
            if (ms.limit_speed) dabs = dabs < ms.min_dabs ? ms.min_dabs : (dabs > ms.max_dabs ? ms.max_dabs : dabs);
            move_hand(a2, st, 2u, dabs, len2, ms.hold_end);
        }
    } else if (st & 1u) {
        if (st & 4u) move_hand(a1, st, 1u, ms.max_speed_dt, len1, ms.hold_end);
    } else if (st & 2u) {
        if (st & 8u) move_hand(a2, st, 2u, ms.max_speed_dt, len2, ms.hold_end);
    }
    aa[k] = a1;
    ab[k] = a2;
    state[k] = (uint8_t)st;
}

void launch_step_couples(Context& c, const MotorStep& ms)
{
    if (!c.ncouples) return;
    const dim3 b(256), g((c.ncouples + 255) / 256);
    if (c.dim == 3)
        hipLaunchKernelGGL(k_step_couples<3>, g, b, 0, c.stream, c.ncouples, c.cpl_fa.get(), c.cpl_aa.get(),
                           c.cpl_fb.get(), c.cpl_ab.get(), c.cpl_w.get(), c.cpl_state.get(), c.d_np.get(),
                           c.d_idx.get(), c.d_cut.get(), c.d_pos.get(), ms);
    else
        hipLaunchKernelGGL(k_step_couples<2>, g, b, 0, c.stream, c.ncouples, c.cpl_fa.get(), c.cpl_aa.get(),
                           c.cpl_fb.get(), c.cpl_ab.get(), c.cpl_w.get(), c.cpl_state.get(), c.d_np.get(),
                           c.d_idx.get(), c.d_cut.get(), c.d_pos.get(), ms);
    CYMG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Meca::solve, first part (meca.cc:621-693): vFOR = vBAS + (R + B + C) vPTS,
// Brownian terms b*N(0,1) with b = sqrt(2 kT/dt / mobility) (rigid_fiber.cc:90-98),
// vRHS = dt * mobility * P vFOR keeping the multipliers in rfLag
// (rigid_fiber.cc:124-132), P' coefficients (rigid_fiber_project.cc:168-219).
// Per-fiber partials of noise = min(mobility*b) and bnorm = |vRHS|_inf; the
// solver's first kernel reduces them per system.
template <int D, int P>
__global__ __launch_bounds__(kWave) void k_setup(FiberArgs a, const double* __restrict__ pts,
                                                 const double* __restrict__ bas, const double* __restrict__ gauss,
                                                 double kT_dt, double* __restrict__ vfor, double* __restrict__ rhs,
                                                 double* __restrict__ lag, double* __restrict__ pdiff,
                                                 double* __restrict__ part, uint32_t nfib)
{
    __shared__ FiberLds<D, P> s;
    const uint32_t f = blockIdx.x;
    const uint32_t np = a.np[f], idx = a.idx[f], seg = a.seg[f];
    const int S = (int)np - 1;
    const uint32_t o = D * idx;
    for (int i = threadIdx.x; i < (int)(D * np); i += kWave) s.x[i] = pts[o + i];
    load_factors<D>(a, s, seg, S);
    __syncthreads();
    fiber_forces<D>(a, s, pts, np, idx, seg, a.rfrig[f], true, false);
    __syncthreads();
    const double mobf = a.mob[f];
    const double b = sqrt(2 * kT_dt / mobf);
    for (int i = threadIdx.x; i < (int)(D * np); i += kWave) {
        double v = bas[o + i] + s.f[i];
        if (gauss) v += b * gauss[o + i];
        s.f[i] = v;
        vfor[o + i] = v;
    }
    __syncthreads();
    fiber_project<D>(a, s, np, seg, a.dt * mobf);
    __syncthreads();
    double bmax = 0.0;
    for (int i = threadIdx.x; i < (int)(D * np); i += kWave) {
        rhs[o + i] = s.y[i];
        bmax = nan_max(bmax, fabs(s.y[i]));
    }
    const double isc = 1.0 / a.cut[f];
    for (int q = threadIdx.x; q < S; q += kWave) {
        const double l = s.t[q];
        lag[seg + q] = l;
        pdiff[seg + q] = (l < 0) ? 0.0 : l * isc;
    }
    bmax = wave_max(bmax);
    if (threadIdx.x == 0) {
        part[P_NOISE * nfib + f] = mobf * b;
        part[P_BNORM * nfib + f] = bmax;
    }
}

void launch_setup(Context& c, const double* gauss)
{
    launch_solid_setup(c, gauss);
    if (!c.nfib) return;
    FiberArgs a = make_fiber_args(c);
    const dim3 g(c.nfib), b(kWave);
    const double kT_dt = c.prm.kT / c.prm.time_step;
    if (c.dim == 3)
        CYMG_FIBER_LAUNCH(k_setup, 3, g, b, 0, c.stream, a, c.d_pts.get(), c.d_bas.get(), gauss, kT_dt,
                           c.d_for.get(), c.d_rhs.get(), c.d_lag.get(), c.d_pdiff.get(), c.d_part.get(),
                           c.nmec());
    else
        CYMG_FIBER_LAUNCH(k_setup, 2, g, b, 0, c.stream, a, c.d_pts.get(), c.d_bas.get(), gauss, kT_dt,
                           c.d_for.get(), c.d_rhs.get(), c.d_lag.get(), c.d_pdiff.get(), c.d_part.get(),
                           c.nmec());
    CYMG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Meca::multiply (meca.cc:157-181) as a standalone launch
template <int D, int P>
__global__ __launch_bounds__(kWave) void k_matvec(FiberArgs a, const double* __restrict__ X, double* __restrict__ Y)
{
    __shared__ FiberLds<D, P> s;
    const uint32_t f = blockIdx.x;
    const uint32_t np = a.np[f], o = D * a.idx[f];
    for (int i = threadIdx.x; i < (int)(D * np); i += kWave) s.x[i] = X[o + i];
    load_factors<D>(a, s, a.seg[f], (int)np - 1);
    __syncthreads();
    fiber_operator<D>(a, s, X, f);
    for (int i = threadIdx.x; i < (int)(D * np); i += kWave) Y[o + i] = s.y[i];
}

void launch_matvec(Context& c, const double* X, double* Y)
{
    launch_solid_matvec(c, X, Y);
    if (!c.nfib) return;
    FiberArgs a = make_fiber_args(c);
    const dim3 g(c.nfib), b(kWave);
    if (c.dim == 3)
        CYMG_FIBER_LAUNCH(k_matvec, 3, g, b, 0, c.stream, a, X, Y);
    else
        CYMG_FIBER_LAUNCH(k_matvec, 2, g, b, 0, c.stream, a, X, Y);
    CYMG_HIP(hipGetLastError());
}

// Meca::computeForces(X, Y, with_rigidity) (meca.cc:143-147): Y = vBAS + (R) + (B + C) X
template <int D, int P>
__global__ __launch_bounds__(kWave) void k_forces(FiberArgs a, const double* __restrict__ X,
                                                  const double* __restrict__ bas, double* __restrict__ Y, int rig)
{
    __shared__ FiberLds<D, P> s;
    const uint32_t f = blockIdx.x;
    const uint32_t np = a.np[f], idx = a.idx[f], o = D * idx;
    for (int i = threadIdx.x; i < (int)(D * np); i += kWave) s.x[i] = X[o + i];
    __syncthreads();
    fiber_forces<D>(a, s, X, np, idx, a.seg[f], a.rfrig[f], rig != 0, false);
    __syncthreads();
    for (int i = threadIdx.x; i < (int)(D * np); i += kWave) Y[o + i] = bas[o + i] + s.f[i];
}

void launch_forces(Context& c, const double* X, double* Y, bool with_rigidity)
{
    launch_solid_forces(c, X, Y);
    if (!c.nfib) return;
    FiberArgs a = make_fiber_args(c);
    const dim3 g(c.nfib), b(kWave);
    if (c.dim == 3)
        CYMG_FIBER_LAUNCH(k_forces, 3, g, b, 0, c.stream, a, X, c.d_bas.get(), Y, with_rigidity ? 1 : 0);
    else
        CYMG_FIBER_LAUNCH(k_forces, 2, g, b, 0, c.stream, a, X, c.d_bas.get(), Y, with_rigidity ? 1 : 0);
    CYMG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// post-solve: vPTS += vSOL (meca.cc:802)
__global__ void k_axpy(uint32_t n, const double* __restrict__ x, double* __restrict__ y)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = y[i] + x[i];
}

// FiberNaked::getPoints (fiber_naked.cc:414-425): reshape_two (172-193),
// reshape_it (221-354: Newton on the segment lengths, dgtsv solve, stop at
// sum|dsca| <= 0.001 or after 32 iterations), reshape_sure (367-401) on failure.
template <int D, int P>
__global__ __launch_bounds__(kWave) void k_reshape(const uint32_t* __restrict__ npv, const uint32_t* __restrict__ idxv,
                                                   const double* __restrict__ cutv, const double* __restrict__ src,
                                                   double* __restrict__ dst)
{
    __shared__ double dif[3 * P], vec[3 * P];
    __shared__ double sca[P], val[P], dia[P];
    __shared__ double low[P], upe[P];
    __shared__ int fail;
    const uint32_t f = blockIdx.x;
    const uint32_t np = npv[f], o = D * idxv[f];
    const int ns = (int)np - 1;
    const double cut = cutv[f];
    const double* X = src + o;
    double* Y = dst + o;
    if (ns == 1) {
        if (threadIdx.x == 0) {
            double df[D], nn = 0.0;
            for (int d = 0; d < D; ++d) { df[d] = X[D + d] - X[d]; nn += df[d] * df[d]; }
            const double sc = 0.5 * (1 - cut / sqrt(nn));
            for (int d = 0; d < D; ++d) {
                const double v = df[d] * sc;
                Y[d] = X[d] + v;
                Y[D + d] = X[D + d] - v;
            }
        }
        return;
    }
    const double a2 = cut * cut;
    if (ns <= kWave) {
        // Newton on the wave: lane p owns segment p, neighbours through shuffles,
        // the tridiagonal Jacobian solved by cyclic reduction (the dgtsv system;
        // no row interchanges are needed, the Jacobian is diagonally dominant
        // near the constraint)
        const int p = threadIdx.x;
        const bool on = p < ns;
        double df[D];
#pragma unroll
        for (int d = 0; d < D; ++d) df[d] = on ? X[D * p + D + d] - X[D * p + d] : 0.0;
        double sc = 0.0;
        int ok = 0;
        for (int it = 0; it <= 32; ++it) {
            const double scm = __shfl_up(sc, 1, kWave), scp = __shfl_down(sc, 1, kWave);
            double dm[D], dp[D], vec[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                dm[d] = __shfl_up(df[d], 1, kWave);
                dp[d] = __shfl_down(df[d], 1, kWave);
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double v = (1 - 2 * sc) * df[d];
                if (p > 0) v = scm * dm[d] + v;
                if (p + 1 < ns) v = v + scp * dp[d];
                vec[d] = v;
            }
            double vv = 0, vd = 0, vl = 0, vu = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                vv += vec[d] * vec[d];
                vd += vec[d] * df[d];
                vl += vec[d] * dm[d];
                vu += vec[d] * dp[d];
            }
            const double val = vv - a2;
            const double dia = -2 * vd;
            const double low = (p > 0) ? vl : 0.0;
            const double upe = (p + 1 < ns) ? vu : 0.0;
            const double x = pcr_solve(ns, low, on ? dia : 1.0, upe, on ? val : 0.0);
            sc += on ? -0.5 * x : 0.0;
            double err = wave_sum(on ? fabs(x) : 0.0);
            if (!(err == err) || err > 1e300) break;        // singular system: reshape_sure
            if (it + 1 > 32) break;                           // iteration cap of reshape_it
            if (!(err > 0.001)) { ok = 1; break; }
        }
        if (ok) {
            double e[D], dd[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                dd[d] = sc * df[d];
                e[d] = __shfl_up(dd[d], 1, kWave);
            }
            if (on) {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    double v = X[D * p + d] + dd[d];
                    if (p > 0) v = v - e[d];
                    Y[D * p + d] = v;
                }
                if (p == ns - 1)
#pragma unroll
                    for (int d = 0; d < D; ++d) Y[D * ns + d] = X[D * ns + d] - dd[d];
            }
            return;
        }
        if (threadIdx.x == 0) fail = 1;
        __syncthreads();
    } else {
    for (int p = threadIdx.x; p < ns; p += kWave) {
        for (int d = 0; d < 3; ++d) dif[3 * p + d] = d < D ? X[D * p + D + d] - X[D * p + d] : 0.0;
        sca[p] = 0.0;
    }
    if (threadIdx.x == 0) fail = 0;
    __syncthreads();
    int it = 0;
    while (true) {
        for (int p = threadIdx.x; p < ns; p += kWave) {
            for (int d = 0; d < 3; ++d) {
                double v;
                if (p == 0) v = (1 - 2 * sca[0]) * dif[d] + sca[1] * dif[3 + d];
                else if (p == ns - 1) v = sca[ns - 2] * dif[3 * (ns - 2) + d] + (1 - 2 * sca[ns - 1]) * dif[3 * (ns - 1) + d];
                else v = sca[p - 1] * dif[3 * (p - 1) + d] + (1 - 2 * sca[p]) * dif[3 * p + d] + sca[p + 1] * dif[3 * (p + 1) + d];
                vec[3 * p + d] = v;
            }
        }
        __syncthreads();
        for (int p = threadIdx.x; p < ns; p += kWave) {
            double vv = 0, vd = 0;
            for (int d = 0; d < 3; ++d) {
                vv += vec[3 * p + d] * vec[3 * p + d];
                vd += vec[3 * p + d] * dif[3 * p + d];
            }
            val[p] = vv - a2;
            dia[p] = -2 * vd;
            if (p >= 1) {
                double vl = 0, vu = 0;
                for (int d = 0; d < 3; ++d) {
                    vl += vec[3 * p + d] * dif[3 * (p - 1) + d];
                    vu += vec[3 * (p - 1) + d] * dif[3 * p + d];
                }
                low[p] = vl;
                upe[p - 1] = vu;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // dgtsv(ns, low+1, dia, upe, val): partial pivoting on the tridiagonal
            double* dl = low + 1;
            double* dd = dia;
            double* du = upe;
            double* bb = val;
            int info = 0;
            for (int i = 0; i + 2 < ns && !info; ++i) {
                if (fabs(dd[i]) >= fabs(dl[i])) {
                    if (dd[i] == 0.0) { info = i + 1; break; }
                    const double fct = dl[i] / dd[i];
                    dd[i + 1] -= fct * du[i];
                    bb[i + 1] -= fct * bb[i];
                    dl[i] = 0.0;
                } else {
                    const double fct = dd[i] / dl[i];
                    dd[i] = dl[i];
                    double t = dd[i + 1];
                    dd[i + 1] = du[i] - fct * t;
                    dl[i] = du[i + 1];
                    du[i + 1] = -fct * dl[i];
                    du[i] = t;
                    t = bb[i];
                    bb[i] = bb[i + 1];
                    bb[i + 1] = t - fct * bb[i + 1];
                }
            }
            if (!info) {
                const int i = ns - 2;
                if (fabs(dd[i]) >= fabs(dl[i])) {
                    if (dd[i] == 0.0) info = i + 1;
                    else {
                        const double fct = dl[i] / dd[i];
                        dd[i + 1] -= fct * du[i];
                        bb[i + 1] -= fct * bb[i];
                    }
                } else {
                    const double fct = dd[i] / dl[i];
                    dd[i] = dl[i];
                    double t = dd[i + 1];
                    dd[i + 1] = du[i] - fct * t;
                    du[i] = t;
                    t = bb[i];
                    bb[i] = bb[i + 1];
                    bb[i + 1] = t - fct * bb[i + 1];
                }
            }
            if (!info && dd[ns - 1] == 0.0) info = ns;
            if (!info) {
                bb[ns - 1] /= dd[ns - 1];
                bb[ns - 2] = (bb[ns - 2] - du[ns - 2] * bb[ns - 1]) / dd[ns - 2];
                for (int i = ns - 3; i >= 0; --i) bb[i] = (bb[i] - du[i] * bb[i + 1] - dl[i] * bb[i + 2]) / dd[i];
                double err = 0;
                for (int p = 0; p < ns; ++p) {
                    sca[p] += -0.5 * bb[p];
                    err += fabs(bb[p]);
                }
                ++it;
                if (it > 32) fail = 1;
                else if (!(err > 0.001)) fail = -1;   // converged
            } else {
                fail = 1;
            }
        }
        __syncthreads();
        if (fail) break;
    }
    if (fail < 0) {
        if (threadIdx.x == 0) {
            double e[3], dd[3] = { 0, 0, 0 };
            for (int d = 0; d < 3; ++d) e[d] = sca[0] * dif[d];
            for (int d = 0; d < D; ++d) Y[d] = X[d] + e[d];
            for (int p = 1; p < ns; ++p) {
                for (int d = 0; d < 3; ++d) dd[d] = sca[p] * dif[3 * p + d];
                for (int d = 0; d < D; ++d) Y[D * p + d] = X[D * p + d] + dd[d] - e[d];
                for (int d = 0; d < 3; ++d) e[d] = dd[d];
            }
            for (int d = 0; d < D; ++d) Y[D * ns + d] = X[D * ns + d] - dd[d];
        }
        return;
    }
    }
    // reshape_sure on the unmodified points
    if (threadIdx.x == 0) {
        for (int i = 0; i < D * (ns + 1); ++i) Y[i] = X[i];
        double dp[3] = { 0, 0, 0 }, sum[3] = { 0, 0, 0 }, sg[3];
        double dis = 0;
        for (int d = 0; d < D; ++d) { sg[d] = Y[D + d] - Y[d]; dis += sg[d] * sg[d]; }
        dis = sqrt(dis);
        if (dis > kRealEpsilon)
            for (int d = 0; d < D; ++d) dp[d] = (cut / dis - 1.0) * sg[d];
        for (int p = 1; p < ns; ++p) {
            dis = 0;
            for (int d = 0; d < D; ++d) { sg[d] = Y[D * p + D + d] - Y[D * p + d]; dis += sg[d] * sg[d]; }
            dis = sqrt(dis);
            for (int d = 0; d < D; ++d) { Y[D * p + d] += dp[d]; sum[d] += dp[d]; }
            if (dis > kRealEpsilon)
                for (int d = 0; d < D; ++d) dp[d] += (cut / dis - 1.0) * sg[d];
        }
        for (int d = 0; d < D; ++d) Y[D * ns + d] += dp[d];
        for (int d = 0; d < D; ++d) sum[d] = (sum[d] + dp[d]) * (-1.0 / (ns + 1));
        for (int p = 0; p <= ns; ++p)
            for (int d = 0; d < D; ++d) Y[D * p + d] += sum[d];
    }
}

void launch_apply(Context& c)
{
    const uint32_t n = c.n();
    if (!n) return;
    hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, c.d_sol.get(), c.d_pts.get());
    CYMG_HIP(hipGetLastError());
    launch_forces(c, c.d_pts.get(), c.d_for.get(), false);   // meca.cc:826
    launch_solid_points(c);
    if (!c.nfib) return;
    const dim3 g(c.nfib), b(kWave);
    if (c.dim == 3)
        CYMG_FIBER_LAUNCH(k_reshape, 3, g, b, 0, c.stream, c.d_np.get(), c.d_idx.get(), c.d_cut.get(), c.d_pts.get(),
                           c.d_pos.get());
    else
        CYMG_FIBER_LAUNCH(k_reshape, 2, g, b, 0, c.stream, c.d_np.get(), c.d_idx.get(), c.d_cut.get(), c.d_pts.get(),
                           c.d_pos.get());
    CYMG_HIP(hipGetLastError());
}

// The diagonal block of fiber f (Meca::getBlock, meca.cc:231-283) for blocks
// too large for the workgroup-local builder: column j is the fiber's operator
// (R + B + C + P', projection, identity) applied to the unit vector e_j, with
// the couplings to other Mecables reading zero.  One wavefront per (fiber,
// column); columns j = blockIdx.y, blockIdx.y + gridDim.y, ...; M column-major.
struct UnitOperand {
    uint32_t at;   // the one non-zero entry of the global operand
    __device__ __forceinline__ double operator()(uint32_t j) const { return j == at ? 1.0 : 0.0; }
};

template <int D>
__global__ __launch_bounds__(kWave) void k_block_columns(FiberArgs a, const uint32_t* __restrict__ list,
                                                         const uint64_t* __restrict__ off, int by_position,
                                                         double* __restrict__ M)
{
    __shared__ FiberLds<D> s;
    const uint32_t f = list[blockIdx.x];
    const uint32_t np = a.np[f], o = D * a.idx[f];
    const int bs = D * (int)np;
    double* B = M + (by_position ? off[blockIdx.x] : off[f]);
    load_factors<D>(a, s, a.seg[f], (int)np - 1);
    for (int j = blockIdx.y; j < bs; j += gridDim.y) {
        for (int i = threadIdx.x; i < bs; i += kWave) s.x[i] = (i == j) ? 1.0 : 0.0;
        __syncthreads();
        fiber_operator_op<D>(a, s, UnitOperand{ o + (uint32_t)j }, f);
        for (int i = threadIdx.x; i < bs; i += kWave) B[(size_t)bs * j + i] = s.y[i];
        __syncthreads();
    }
}

// off: block offsets by fiber (by_position 0) or by position in list
void launch_block_columns(Context& c, const uint32_t* list, uint32_t count, const uint64_t* off, int by_position,
                          double* M)
{
    if (!count) return;
    FiberArgs a = make_fiber_args(c);
    // enough (fiber, column) wavefronts to fill the device several times
    const uint32_t cols = std::max<uint32_t>(1, std::min<uint32_t>(c.dim * c.largest, 4096u / count + 1));
    const dim3 g(count, cols), b(kWave);
    if (c.dim == 3)
        hipLaunchKernelGGL(k_block_columns<3>, g, b, 0, c.stream, a, list, off, by_position, M);
    else
        hipLaunchKernelGGL(k_block_columns<2>, g, b, 0, c.stream, a, list, off, by_position, M);
    CYMG_HIP(hipGetLastError());
}

// RigidFiber::computeTensions (rigid_fiber_project.cc:153-161) on vFOR
template <int D, int P>
__global__ __launch_bounds__(kWave) void k_tensions(FiberArgs a, const double* __restrict__ F, double* __restrict__ lag)
{
    __shared__ FiberLds<D, P> s;
    const uint32_t f = blockIdx.x;
    const uint32_t np = a.np[f], o = D * a.idx[f], seg = a.seg[f];
    const int S = (int)np - 1;
    for (int i = threadIdx.x; i < (int)(D * np); i += kWave) s.f[i] = F[o + i];
    load_factors<D>(a, s, seg, S, true);
    __syncthreads();
    for (int q = threadIdx.x; q < S; q += kWave) {
        double t = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) t += a.diff[D * (seg + q) + d] * (s.f[D * (q + 1) + d] - s.f[D * q + d]);
        s.t[q] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) ptts2_serial(S, s.d, s.e, s.t);
    __syncthreads();
    for (int q = threadIdx.x; q < S; q += kWave) lag[seg + q] = s.t[q];
}

void launch_tensions(Context& c)
{
    if (!c.nfib) return;
    FiberArgs a = make_fiber_args(c);
    const dim3 g(c.nfib), b(kWave);
    if (c.dim == 3)
        CYMG_FIBER_LAUNCH(k_tensions, 3, g, b, 0, c.stream, a, c.d_for.get(), c.d_lag.get());
    else
        CYMG_FIBER_LAUNCH(k_tensions, 2, g, b, 0, c.stream, a, c.d_for.get(), c.d_lag.get());
    CYMG_HIP(hipGetLastError());
}

}  // namespace cymg
