// Kernels of the Solid Mecable (solid.cc) outside the Krylov iteration:
// prepare (drag, centre, momentum factor), setup (forces, Brownian terms,
// right-hand side), the standalone operator and forces, the preconditioner
// block in low-rank form and its application, and the post-solve getPoints
// (rescale, or reshape every 8th call).  One workgroup per Solid.
#include "args.h"

namespace cymg {

// Solid::prepareMecable (solid.cc:707-711): setDragCoefficient (665-704),
// makeProjection (2D: 757-760, 3D: 789-847; soCenter = centroid(), 181-200),
// momentum matrix factor by dpotf2('U'); putPoints into vPTS (meca.cc:488-537).
template <int D>
__global__ __launch_bounds__(kSolidThreads) void k_solid_prepare(SolidArgs s, const double* __restrict__ pos,
                                                                 const double* __restrict__ shape_sqr,
                                                                 double* __restrict__ pts, double* __restrict__ geo,
                                                                 int* __restrict__ status, double visc)
{
    __shared__ double red[16 * 12];
    const uint32_t k = blockIdx.x;
    const uint32_t np = s.np[k], idx = s.idx[k];
    for (uint32_t i = threadIdx.x; i < D * np; i += blockDim.x) pts[D * idx + i] = pos[D * idx + i];
    // 0 sum r, 1 sum r^3, 2..4 sum r x, 5 sum r |x|^2 (2D), 6..11 sum r x_i x_j (3D)
    double v[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) v[q] = 0.0;
    for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
        const double r = s.rad[idx + p - s.p0];
        if (r > 0) {
            const double* x = pos + D * (idx + p);
            v[0] += r;
            v[1] += r * r * r;
#pragma unroll
            for (int d = 0; d < D; ++d) v[2 + d] += r * x[d];
            if constexpr (D == 2) v[5] += r * (x[0] * x[0] + x[1] * x[1]);
            if constexpr (D == 3) {
                const double px = r * x[0], py = r * x[1], pz = r * x[2];
                v[6] += px * x[0]; v[7] += px * x[1]; v[8] += px * x[2];
                v[9] += py * x[1]; v[10] += py * x[2]; v[11] += pz * x[2];
            }
        }
    }
    wg_sum<12>(v, red);
    if (threadIdx.x == 0) {
        double* g = geo + kSolidGeo * k;
        const double eta6 = 6 * M_PI * visc;
        int bad = 0;
        if (v[0] < 128.0 * 2.220446049250313e-16) bad = 1;    // "zero drag in Solid"
        double c[3] = { 0, 0, 0 };
        for (int d = 0; d < D; ++d) c[d] = v[2 + d] / v[0];
        const double drag = v[0] * eta6, rot = v[1] * (8 * M_PI * visc);
        g[SG_DRAG] = drag;
        g[SG_ROT] = rot;
        g[SG_SHAPE_SQR] = shape_sqr[k];
        for (int d = 0; d < 3; ++d) g[SG_C + d] = c[d];
        if constexpr (D == 2) {
            const double m2 = rot + eta6 * v[5] - drag * (c[0] * c[0] + c[1] * c[1]);
            g[SG_MOM2D] = m2;
            if (!(m2 >= 128.0 * 2.220446049250313e-16)) bad = 1;
        } else {
            const double m0 = v[6] * eta6, m3 = v[7] * eta6, m6 = v[8] * eta6;
            const double m4 = v[9] * eta6, m7 = v[10] * eta6, m8 = v[11] * eta6;
            const double diag = m0 + m4 + m8 + rot - drag * (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            double A[9] = { 0 };
            A[0 + 3 * 0] = diag - m0 + drag * c[0] * c[0];
            A[0 + 3 * 1] = -m3 + drag * c[0] * c[1];
            A[0 + 3 * 2] = -m6 + drag * c[0] * c[2];
            A[1 + 3 * 1] = diag - m4 + drag * c[1] * c[1];
            A[1 + 3 * 2] = -m7 + drag * c[1] * c[2];
            A[2 + 3 * 2] = diag - m8 + drag * c[2] * c[2];
            // dpotf2('U')
            for (int j = 0; j < 3 && !bad; ++j) {
                double ajj = A[j + 3 * j];
                for (int q = 0; q < j; ++q) ajj -= A[q + 3 * j] * A[q + 3 * j];
                if (!(ajj > 0.0)) { bad = 1; break; }
                ajj = sqrt(ajj);
                A[j + 3 * j] = ajj;
                for (int i = j + 1; i < 3; ++i) {
                    double t = A[j + 3 * i];
                    for (int q = 0; q < j; ++q) t -= A[q + 3 * j] * A[q + 3 * i];
                    A[j + 3 * i] = t / ajj;
                }
            }
            double* U = g + SG_U;
            U[0] = A[0]; U[1] = A[3]; U[2] = A[6]; U[3] = A[4]; U[4] = A[7]; U[5] = A[8];
        }
        if (bad) atomicOr(status, 2);
    }
}

// Meca::solve setup rows of Solid k (meca.cc:622-651): vFOR = vBAS + (B + C) vPTS
// + b N(0,1) with b = sqrt(2 (kT/dt) soDrag / np) (Solid::addBrownianForces,
// solid.cc:714-725, noise b / soDrag); vRHS = setSpeedsFromForces(vFOR, dt)
template <int D>
__global__ __launch_bounds__(kSolidThreads) void k_solid_setup(FiberArgs a, SolidArgs s, const double* __restrict__ pts,
                                                               const double* __restrict__ bas,
                                                               const double* __restrict__ gauss, double kT_dt,
                                                               double* __restrict__ vfor, double* __restrict__ rhs,
                                                               double* __restrict__ part, uint32_t stride)
{
    constexpr int NG = RigidDim<D>::NG;
    __shared__ double red[16 * 6];
    const uint32_t k = blockIdx.x;
    const uint32_t np = s.np[k], idx = s.idx[k];
    const double* geo = s.geo + kSolidGeo * k;
    const double b = sqrt(2 * kT_dt * geo[SG_DRAG] / np);
    double acc[6] = { 0, 0, 0, 0, 0, 0 };
    for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
        double fp[D];
        coupling_row<D>(a, idx + p, Plain{ pts }, fp);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const uint32_t i = D * (idx + p) + d;
            double v = bas[i] + fp[d];
            if (gauss) v += b * gauss[i];
            fp[d] = v;
            vfor[i] = v;
        }
        add_moments<D>(pts + D * (idx + p), fp, acc);
    }
    wg_sum<NG>(acc, red);
    double q[6];
    solid_q<D>(geo, a.dt, acc, q);
    double bmax = 0.0;
    for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
        const double* x = pts + D * (idx + p);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double v = rigid_velocity<D>(q, x, d);
            rhs[D * (idx + p) + d] = v;
            bmax = nan_max(bmax, fabs(v));
        }
    }
    bmax = wg_max(bmax, red);
    if (threadIdx.x == 0) {
        part[P_NOISE * stride + s.nfib + k] = b / geo[SG_DRAG];
        part[P_BNORM * stride + s.nfib + k] = bmax;
    }
}

constexpr uint32_t kSolidRows = 4096;   // row extents staged in LDS by the Solid kernels

struct StoreTo {
    double* __restrict__ Y;
    __device__ __forceinline__ void operator()(uint32_t i, double v) const { Y[i] = v; }
};

template <int D>
__global__ __launch_bounds__(kSolidThreads) void k_solid_matvec(FiberArgs a, SolidArgs s, const double* __restrict__ X,
                                                                double* __restrict__ Y)
{
    __shared__ double red[16 * 6];
    __shared__ uint32_t lrow[kSolidRows];
    solid_operator<D>(a, s, blockIdx.x, Plain{ X }, StoreTo{ Y }, red, lrow, kSolidRows);
}

// Meca::computeForces(X, Y) rows of Solid k: Y = vBAS + (B + C) X (a Solid has no rigidity)
template <int D>
__global__ __launch_bounds__(kSolidThreads) void k_solid_forces(FiberArgs a, SolidArgs s, const double* __restrict__ X,
                                                                const double* __restrict__ bas, double* __restrict__ Y)
{
    const uint32_t k = blockIdx.x;
    const uint32_t np = s.np[k], idx = s.idx[k];
    for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
        double fp[D];
        coupling_row<D>(a, idx + p, Plain{ X }, fp);
#pragma unroll
        for (int d = 0; d < D; ++d) Y[D * (idx + p) + d] = bas[D * (idx + p) + d] + fp[d];
    }
}

// Preconditioner block of Solid k (Meca::getBlock, meca.cc:231-283, inverted
// as in computePreconditionner, meca.cc:323-444) in the low-rank form:
// H = K_ss G (the Solid's columns of B (x) I + C applied to the rigid basis),
// A = -dt Q1 H', S = I + A G, W = S^-1 A.  A is kept (for the block export),
// W is what the preconditioner reads.  useW = 0 when S is singular (the
// reference falls back to the identity, meca.cc:437-438).
template <int D>
__global__ __launch_bounds__(kSolidThreads) void k_solid_pc_build(FiberArgs a, SolidArgs s, double* __restrict__ Aout,
                                                                  double* __restrict__ Wout, int* __restrict__ useW)
{
    constexpr int NG = RigidDim<D>::NG;
    __shared__ double red[16 * 36];
    __shared__ double Sinv[36];
    __shared__ int ok;
    const uint32_t k = blockIdx.x;
    const uint32_t np = s.np[k], idx = s.idx[k];
    const uint32_t o = D * idx, n = D * np;
    const double* geo = s.geo + kSolidGeo * k;
    double* A = Aout + (size_t)6 * (o - D * s.p0);
    double* W = Wout + (size_t)6 * (o - D * s.p0);
    double ag[NG * NG];
#pragma unroll
    for (int q = 0; q < NG * NG; ++q) ag[q] = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t p = i / D;
        const int d = (int)(i - D * p);
        const uint32_t r = idx + p;
        double h[6] = { 0, 0, 0, 0, 0, 0 };
        double g[6];
        if (a.Buse) {
            // the columns of a long row (an aster centre) are loaded 8 at a time
            const uint32_t e1 = a.Brow[r + 1];
            uint32_t q = a.Brow[r];
            for (; q + 8 <= e1; q += 8) {
                uint32_t cq[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) cq[j] = a.Bcol[q + j];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t cp = cq[j];
                    if (cp < idx || cp >= idx + np) continue;
                    rigid_row<D>(s.pos + D * cp, d, g);
                    const double v = a.Bval[q + j];
#pragma unroll
                    for (int c = 0; c < NG; ++c) h[c] += v * g[c];
                }
            }
            for (; q < e1; ++q) {
                const uint32_t cp = a.Bcol[q];
                if (cp < idx || cp >= idx + np) continue;
                rigid_row<D>(s.pos + D * cp, d, g);
                const double v = a.Bval[q];
#pragma unroll
                for (int c = 0; c < NG; ++c) h[c] += v * g[c];
            }
        }
        if (a.Cuse) {
            const uint32_t e1 = a.Crow[D * r + d + 1];
            for (uint32_t q = a.Crow[D * r + d]; q < e1; ++q) {
                const uint32_t j = a.Ccol[q];
                const uint32_t cp = j / D;
                if (cp < idx || cp >= idx + np) continue;
                rigid_row<D>(s.pos + D * cp, (int)(j - D * cp), g);
                const double v = a.Cval[q];
#pragma unroll
                for (int c = 0; c < NG; ++c) h[c] += v * g[c];
            }
        }
        double av[6];
        solid_q<D>(geo, -a.dt, h, av);
#pragma unroll
        for (int c = 0; c < 6; ++c) A[6 * i + c] = c < NG ? av[c] : 0.0;
        rigid_row<D>(s.pos + D * r, d, g);
#pragma unroll
        for (int c = 0; c < NG; ++c)
#pragma unroll
            for (int e = 0; e < NG; ++e) ag[c * NG + e] += av[c] * g[e];
    }
    wg_sum<NG * NG>(ag, red);
    if (threadIdx.x == 0) {
        // Gauss-Jordan with partial pivoting on S = I + A G
        double M[NG][2 * NG];
        for (int c = 0; c < NG; ++c)
            for (int e = 0; e < NG; ++e) {
                M[c][e] = (c == e ? 1.0 : 0.0) + ag[c * NG + e];
                M[c][NG + e] = c == e ? 1.0 : 0.0;
            }
        int good = 1;
        for (int col = 0; col < NG && good; ++col) {
            int piv = col;
            for (int r = col + 1; r < NG; ++r)
                if (fabs(M[r][col]) > fabs(M[piv][col])) piv = r;
            if (!(fabs(M[piv][col]) > 0.0)) { good = 0; break; }
            if (piv != col)
                for (int e = 0; e < 2 * NG; ++e) { const double t = M[col][e]; M[col][e] = M[piv][e]; M[piv][e] = t; }
            const double inv = 1.0 / M[col][col];
            for (int e = 0; e < 2 * NG; ++e) M[col][e] *= inv;
            for (int r = 0; r < NG; ++r) {
                if (r == col) continue;
                const double m = M[r][col];
                if (m != 0.0)
                    for (int e = 0; e < 2 * NG; ++e) M[r][e] -= m * M[col][e];
            }
        }
        for (int c = 0; c < NG; ++c)
            for (int e = 0; e < NG; ++e) {
                const double v = M[c][NG + e];
                if (!(fabs(v) < __builtin_inf())) good = 0;
                Sinv[c * NG + e] = v;
            }
        ok = good;
        useW[k] = good;
    }
    __syncthreads();
    if (!ok) return;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        double av[NG];
#pragma unroll
        for (int c = 0; c < NG; ++c) av[c] = A[6 * i + c];   // written by this thread above
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double w = 0.0;
            if (c < NG)
#pragma unroll
                for (int e = 0; e < NG; ++e) w += Sinv[c * NG + e] * av[e];
            W[6 * i + c] = w;
        }
    }
}

struct ReadFrom {
    const double* __restrict__ X;
    __device__ __forceinline__ double operator()(uint32_t i) const { return X[i]; }
};

template <int D>
__global__ __launch_bounds__(kSolidThreads) void k_solid_pc_apply(SolidArgs s, const double* __restrict__ X,
                                                                  double* __restrict__ Y)
{
    __shared__ double red[16 * 6];
    solid_precond<D>(s, blockIdx.x, ReadFrom{ X }, StoreTo{ Y }, red);
}

// top eigenvector of a symmetric 4x4 by cyclic Jacobi rotations (the
// quaternion of Horn's method; LAPACK dsyevx in the reference)
__device__ void sym4_top_eigenvector(double A[16], double v[4])
{
    double V[16];
    for (int i = 0; i < 16; ++i) V[i] = (i % 5 == 0) ? 1.0 : 0.0;
    // cyclic sweeps until the off-diagonal part is below the rounding of the
    // diagonal (Jacobi converges quadratically: a few sweeps); the loops over
    // the pairs are unrolled so that A and V stay in registers
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, dia = 0.0;
#pragma unroll
        for (int q = 1; q < 4; ++q)
#pragma unroll
            for (int p = 0; p < q; ++p) off += A[p + 4 * q] * A[p + 4 * q];
#pragma unroll
        for (int p = 0; p < 4; ++p) dia += A[p + 4 * p] * A[p + 4 * p];
        if (!(off > 1e-34 * dia)) break;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = p + 1; q < 4; ++q) {
                const double apq = A[p + 4 * q];
                if (apq == 0.0) continue;
                const double th = (A[q + 4 * q] - A[p + 4 * p]) / (2 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                const double c = 1 / sqrt(t * t + 1), sn = t * c;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double akp = A[k + 4 * p], akq = A[k + 4 * q];
                    A[k + 4 * p] = c * akp - sn * akq;
                    A[k + 4 * q] = sn * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double apk = A[p + 4 * k], aqk = A[q + 4 * k];
                    A[p + 4 * k] = c * apk - sn * aqk;
                    A[q + 4 * k] = sn * apk + c * aqk;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double vkp = V[k + 4 * p], vkq = V[k + 4 * q];
                    V[k + 4 * p] = c * vkp - sn * vkq;
                    V[k + 4 * q] = sn * vkp + c * vkq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < 4; ++i)
        if (A[i + 4 * i] > A[best + 4 * best]) best = i;
    for (int k = 0; k < 4; ++k) v[k] = V[k + 4 * best];
}

// Solid::getPoints (solid.cc:635-650) after vPTS += vSOL: copy, then
// reshape() on every 8th call (2D: solid.cc:503-540; 3D: 547-620, Horn's
// quaternion, rotation by Quaternion::setMatrix3, quaternion.h:376-400),
// rescale() (solid.cc:434-467) otherwise.  status bit 4: collapsed Solid.
template <int D>
__global__ __launch_bounds__(kSolidThreads) void k_solid_points(SolidArgs s, const double* __restrict__ pts,
                                                                double* __restrict__ pos,
                                                                const double* __restrict__ shape,
                                                                int* __restrict__ timer, int* __restrict__ status)
{
    __shared__ double red[16 * 9];
    __shared__ double rot[9];
    __shared__ double shared_scale;
    const uint32_t k = blockIdx.x;
    const uint32_t np = s.np[k], idx = s.idx[k];
    const double* geo = s.geo + kSolidGeo * k;
    const double* sh = shape + D * (idx - s.p0);
    double m[6] = { 0, 0, 0, 0, 0, 0 };
    for (uint32_t p = threadIdx.x; p < np; p += blockDim.x)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double x = pts[D * (idx + p) + d];
            m[d] += x;
            m[3 + d] += x * x;
        }
    wg_sum<6>(m, red);
    const double N = (double)np;
    double cc[3] = { 0, 0, 0 };
#pragma unroll
    for (int d = 0; d < D; ++d) cc[d] = m[d] * (1.0 / N);
    const int t = timer[k] + 1;
    const bool reshape = t > 7 && np >= 2;
    if (reshape) {
        double S[9] = { 0 };
        for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
            const double* x = pts + D * (idx + p);
            const double* h = sh + D * p;
            if constexpr (D == 3) {
                for (int dd = 0; dd < 3; ++dd)
                    for (int ee = 0; ee < 3; ++ee) S[dd + 3 * ee] += h[dd] * x[ee];
            } else {
                S[0] += x[0] * h[0] + x[1] * h[1];
                S[1] += h[0] * x[1] - h[1] * x[0];
            }
        }
        wg_sum<9>(S, red);
        if (threadIdx.x == 0) {
            if constexpr (D == 3) {
                double mx = fabs(S[0]);
                if (fabs(S[4]) > mx) mx = fabs(S[4]);
                if (fabs(S[8]) > mx) mx = fabs(S[8]);
                const double sc = 1.0 / mx;
                double Nm[16];
                Nm[0 + 4 * 0] = sc * (S[0] + S[4] + S[8]);
                Nm[0 + 4 * 1] = sc * (S[1 + 3 * 2] - S[2 + 3 * 1]);
                Nm[0 + 4 * 2] = sc * (S[2 + 3 * 0] - S[0 + 3 * 2]);
                Nm[0 + 4 * 3] = sc * (S[0 + 3 * 1] - S[1 + 3 * 0]);
                Nm[1 + 4 * 1] = sc * (S[0] - S[4] - S[8]);
                Nm[1 + 4 * 2] = sc * (S[0 + 3 * 1] + S[1 + 3 * 0]);
                Nm[1 + 4 * 3] = sc * (S[2 + 3 * 0] + S[0 + 3 * 2]);
                Nm[2 + 4 * 2] = sc * (S[4] - S[0] - S[8]);
                Nm[2 + 4 * 3] = sc * (S[1 + 3 * 2] + S[2 + 3 * 1]);
                Nm[3 + 4 * 3] = sc * (S[8] - S[4] - S[0]);
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < i; ++j) Nm[i + 4 * j] = Nm[j + 4 * i];
                double q[4];
                sym4_top_eigenvector(Nm, q);
                const double x2 = q[1] + q[1], y2 = q[2] + q[2], z2 = q[3] + q[3];
                const double rx = q[0] * x2, ry = q[0] * y2, rz = q[0] * z2;
                const double xx = q[1] * x2, xy = q[1] * y2, xz = q[1] * z2;
                const double yy = q[2] * y2, yz = q[2] * z2, zz = q[3] * z2;
                rot[0] = 1.0 - (yy + zz); rot[1] = xy + rz; rot[2] = xz - ry;
                rot[3] = xy - rz; rot[4] = 1.0 - (xx + zz); rot[5] = yz + rx;
                rot[6] = xz + ry; rot[7] = yz - rx; rot[8] = 1.0 - (xx + yy);
            } else {
                const double a = S[0], b = S[1];
                const double n = sqrt(a * a + b * b);
                double c = 1, sn = 0;
                if (n > 128.0 * 2.220446049250313e-16) { c = a / n; sn = b / n; }
                rot[0] = c; rot[1] = sn;
            }
            timer[k] = 0;
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
            const double* h = sh + D * p;
            double* y = pos + D * (idx + p);
            if constexpr (D == 3) {
                y[0] = cc[0] + rot[0] * h[0] + rot[3] * h[1] + rot[6] * h[2];
                y[1] = cc[1] + rot[1] * h[0] + rot[4] * h[1] + rot[7] * h[2];
                y[2] = cc[2] + rot[2] * h[0] + rot[5] * h[1] + rot[8] * h[2];
            } else {
                y[0] = rot[0] * h[0] - rot[1] * h[1] + cc[0];
                y[1] = rot[1] * h[0] + rot[0] * h[1] + cc[1];
            }
        }
        return;
    }
    if (threadIdx.x == 0) {
        double sz = 0.0;
        for (int d = 0; d < D; ++d) sz += m[3 + d] - N * cc[d] * cc[d];
        if (np >= 2 && !(sz > 0)) atomicOr(status, 4);
        shared_scale = (np >= 2 && sz > 0) ? sqrt(geo[SG_SHAPE_SQR] / sz) : 1.0;
        timer[k] = np >= 2 ? t : timer[k];
    }
    __syncthreads();
    const double scale = shared_scale;
    for (uint32_t p = threadIdx.x; p < np; p += blockDim.x)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const uint32_t i = D * (idx + p) + d;
            pos[i] = np >= 2 ? scale * (pts[i] - cc[d]) + cc[d] : pts[i];
        }
}

// ---------------------------------------------------------------------------
// launchers

#define SOLID_LAUNCH(kernel, ...)                                                                        \
    do {                                                                                                 \
        const dim3 g_(c.nsol), b_(kSolidThreads);                                                        \
        if (c.dim == 3) hipLaunchKernelGGL(kernel<3>, g_, b_, 0, c.stream, __VA_ARGS__);                  \
        else hipLaunchKernelGGL(kernel<2>, g_, b_, 0, c.stream, __VA_ARGS__);                             \
        CYMG_HIP(hipGetLastError());                                                                     \
    } while (0)

void launch_solid_prepare(Context& c)
{
    if (!c.nsol) return;
    c.d_sol_geo.resize((size_t)kSolidGeo * c.nsol);
    SOLID_LAUNCH(k_solid_prepare, make_solid_args(c), c.d_pos.get(), c.d_sol_shape_sqr.get(), c.d_pts.get(),
                 c.d_sol_geo.get(), c.d_status.get() + 1, c.prm.viscosity);
}

void launch_solid_setup(Context& c, const double* gauss)
{
    if (!c.nsol) return;
    SOLID_LAUNCH(k_solid_setup, make_fiber_args(c), make_solid_args(c), c.d_pts.get(), c.d_bas.get(), gauss,
                 c.prm.kT / c.prm.time_step, c.d_for.get(), c.d_rhs.get(), c.d_part.get(), c.nmec());
}

void launch_solid_matvec(Context& c, const double* X, double* Y)
{
    if (!c.nsol) return;
    SOLID_LAUNCH(k_solid_matvec, make_fiber_args(c), make_solid_args(c), X, Y);
}

void launch_solid_forces(Context& c, const double* X, double* Y)
{
    if (!c.nsol) return;
    SOLID_LAUNCH(k_solid_forces, make_fiber_args(c), make_solid_args(c), X, c.d_bas.get(), Y);
}

void launch_solid_pc_build(Context& c)
{
    if (!c.nsol) return;
    const size_t nd = (size_t)c.dim * (c.npts - c.fib_pts);
    c.d_sol_A.resize(6 * nd);
    c.d_sol_W.resize(6 * nd);
    c.d_sol_useW.resize(c.nsol);
    SOLID_LAUNCH(k_solid_pc_build, make_fiber_args(c), make_solid_args(c), c.d_sol_A.get(), c.d_sol_W.get(),
                 c.d_sol_useW.get());
}

void launch_solid_pc_apply(Context& c, const double* X, double* Y)
{
    if (!c.nsol) return;
    SOLID_LAUNCH(k_solid_pc_apply, make_solid_args(c), X, Y);
}

void launch_solid_points(Context& c)
{
    if (!c.nsol) return;
    SOLID_LAUNCH(k_solid_points, make_solid_args(c), c.d_pts.get(), c.d_pos.get(), c.d_sol_shape.get(),
                 c.d_sol_timer.get(), c.d_status.get() + 2);
}

}  // namespace cymg
