// Block preconditioner of the implicit step (Meca::computePreconditionner,
// meca.cc:323-444, and Meca::precondition, meca.cc:448-466).
//
// One workgroup (4 wavefronts) per fiber builds the dense diagonal block
//   M_f = I - dt*mob*P_f (R_f + B_ff + C_ff + P'_f)        (Meca::getBlock, meca.cc:231-283)
// in LDS, factorises it with partial pivoting (dgetrf semantics: first entry
// of largest magnitude as pivot, zero pivot -> block unused, meca.cc:352-353,437-438),
// inverts it in place (dgetri: inv(U), then X*L = inv(U), then column
// interchanges) and streams the inverse to HBM, column-major, coalesced.
// The FP32 variant builds the block in FP64 and factorises/inverts in FP32.
#include "args.h"
#include "pc_ops.h"

namespace cymg {

constexpr int kPcThreads = 256;

struct PcArgs {
    const uint32_t* np;
    const uint32_t* idx;
    const uint32_t* seg;
    const uint64_t* blkoff;
    const double* mob;
    const double* rfrig;
    const double* diff;
    const double* jd;
    const double* je;
    const double* pdiff;
    const uint32_t* Brow; const uint32_t* Bcol; const double* Bval; int Buse;
    const uint32_t* Crow; const uint32_t* Ccol; const double* Cval; int Cuse;
    double dt;
    int bs_max;
    int factor;   // 0: store the assembled block M itself (for export)
    int method;   // 0: Gauss-Jordan (default), 1: getrf + getri
};

// LU with partial pivoting (dgetf2 order), in place, column-major, lda = bs.
// Returns info through *info (LDS).
template <typename T>
__device__ void lds_getrf(T* A, int bs, int* ipiv, int* info, int* pivrow)
{
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    for (int j = 0; j < bs; ++j) {
        if (tid < kWave) {
            T best = -1;
            int bi = bs;
            for (int i = j + lane; i < bs; i += kWave) {
                const T v = fabs(A[i + bs * j]);
                if (v > best) { best = v; bi = i; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const T ob = __shfl_xor(best, o, kWave);
                const int oi = __shfl_xor(bi, o, kWave);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (lane == 0) {
                if (bi >= bs) bi = j;
                *pivrow = bi;
                ipiv[j] = bi + 1;
            }
        }
        __syncthreads();
        const int p = *pivrow;
        const bool nz = A[p + bs * j] != T(0);
        if (nz) {
            if (p != j)
                for (int k = tid; k < bs; k += kPcThreads) {
                    const T t = A[j + bs * k];
                    A[j + bs * k] = A[p + bs * k];
                    A[p + bs * k] = t;
                }
            __syncthreads();
            const T piv = A[j + bs * j];
            const T r = T(1) / piv;
            for (int i = j + 1 + tid; i < bs; i += kPcThreads) A[i + bs * j] *= r;
        } else if (tid == 0 && *info == 0) {
            *info = j + 1;
        }
        __syncthreads();
        const int m = bs - j - 1;
        for (int t = tid; t < m * m; t += kPcThreads) {
            const int ii = j + 1 + t % m, kk = j + 1 + t / m;
            A[ii + bs * kk] -= A[ii + bs * j] * A[j + bs * kk];
        }
        __syncthreads();
    }
}

// In-place inverse from the LU factors (dgetri order).  work: bs entries.
template <typename T>
__device__ void lds_getri(T* A, int bs, const int* ipiv, T* work)
{
    const int tid = threadIdx.x;
    // inv(U), column by column (dtrti2, upper, non-unit)
    for (int j = 0; j < bs; ++j) {
        if (tid == 0) A[j + bs * j] = T(1) / A[j + bs * j];
        __syncthreads();
        const T ajj = -A[j + bs * j];
        T acc = 0;
        if (tid < j) {
            for (int k = tid; k < j; ++k) acc += A[tid + bs * k] * A[k + bs * j];
        }
        __syncthreads();
        if (tid < j) A[tid + bs * j] = ajj * acc;
        __syncthreads();
    }
    // solve inv(A) * L = inv(U), columns right to left
    for (int j = bs - 1; j >= 0; --j) {
        for (int i = j + 1 + tid; i < bs; i += kPcThreads) {
            work[i] = A[i + bs * j];
            A[i + bs * j] = 0;
        }
        __syncthreads();
        for (int i = tid; i < bs; i += kPcThreads) {
            T acc = A[i + bs * j];
            for (int k = j + 1; k < bs; ++k) acc -= A[i + bs * k] * work[k];
            A[i + bs * j] = acc;
        }
        __syncthreads();
    }
    // undo the row interchanges of the factorisation as column interchanges
    for (int i = tid; i < bs; i += kPcThreads)
        for (int j = bs - 2; j >= 0; --j) {
            const int jp = ipiv[j] - 1;
            if (jp != j) {
                const T t = A[i + bs * j];
                A[i + bs * j] = A[i + bs * jp];
                A[i + bs * jp] = t;
            }
        }
    __syncthreads();
}

// In-place inverse by Gauss-Jordan elimination with partial pivoting (the
// pivot of column k is the first entry of largest magnitude at or below the
// diagonal, as dgetf2 chooses it).  One pass: 3 barriers per column instead of
// the ~9 of getrf + getri, and n^3 multiply-adds instead of ~n^3 spread over
// two triangular phases.  Row interchanges are undone as column
// interchanges at the end (the last step of dgetri).  info != 0: exactly zero
// pivot, the block is not used (meca.cc:437-438).
template <typename T>
__device__ void lds_gauss_jordan(T* A, int bs, int* ipiv, int* info, int* pivrow, T* colk, T* sc)
{
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    for (int k = 0; k < bs; ++k) {
        if (tid < kWave) {
            T best = -1;
            int bi = bs;
            for (int i = k + lane; i < bs; i += kWave) {
                const T v = fabs(A[i + bs * k]);
                if (v > best) { best = v; bi = i; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const T ob = __shfl_xor(best, o, kWave);
                const int oi = __shfl_xor(bi, o, kWave);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (lane == 0) {
                if (bi >= bs) bi = k;
                *pivrow = bi;
                ipiv[k] = bi + 1;
                sc[0] = A[bi + bs * k];   // pivot
                sc[1] = A[k + bs * k];    // entry that moves to row bi
                if (sc[0] == T(0)) *info = k + 1;
            }
        }
        __syncthreads();
        if (*info) return;
        const int p = *pivrow;
        const T r = T(1) / sc[0];
        // multipliers of column k after the interchange (0 on the pivot row)
        for (int i = tid; i < bs; i += kPcThreads) {
            if (i == k) colk[i] = 0;
            else if (i == p) colk[i] = sc[1];
            else colk[i] = A[i + bs * k];
        }
        // interchange rows k and p, scale the pivot row
        for (int j = tid; j < bs; j += kPcThreads) {
            T vk = A[k + bs * j], vp = A[p + bs * j];
            if (p != k) { const T t = vk; vk = vp; vp = t; }
            A[p + bs * j] = vp;
            A[k + bs * j] = (j == k) ? r : vk * r;
        }
        __syncthreads();
        // eliminate column k from every other row
        for (int t = tid; t < bs * bs; t += kPcThreads) {
            const int i = t % bs, j = t / bs;
            if (i == k) continue;
            const T m = colk[i];
            if (j == k) A[i + bs * k] = -m * r;
            else A[i + bs * j] -= m * A[k + bs * j];
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges
    for (int i = tid; i < bs; i += kPcThreads)
        for (int j = bs - 1; j >= 0; --j) {
            const int jp = ipiv[j] - 1;
            if (jp != j) {
                const T t = A[i + bs * j];
                A[i + bs * j] = A[i + bs * jp];
                A[i + bs * jp] = t;
            }
        }
    __syncthreads();
}

template <int D, typename T>
__global__ __launch_bounds__(kPcThreads) void k_precond_build(PcArgs a, T* __restrict__ out, int* __restrict__ useblk)
{
    extern __shared__ double lds[];
    const uint32_t f = blockIdx.x;
    const int np = (int)a.np[f];
    const int S = np - 1;
    const int bs = D * np;
    const uint32_t idx = a.idx[f], seg = a.seg[f];
    const int tid = threadIdx.x;
    // LDS layout
    double* A = lds;                               // bs*bs (built in FP64)
    double* lam = A + (size_t)bs * bs;             // S*bs multipliers, one column per thread
    double* sdif = lam + (size_t)S * bs;           // D*S
    double* sd = sdif + D * S;                     // S
    double* se = sd + S;                           // S
    double* spd = se + S;                          // S
    int* ipiv = (int*)(spd + S);                   // bs
    int* misc = ipiv + bs;                         // info, pivot row
    T* work = (T*)(((uintptr_t)(misc + 4) + 15) & ~(uintptr_t)15);   // bs + 2
    T* AT = (T*)(((uintptr_t)(work + bs + 2) + 15) & ~(uintptr_t)15);   // bs*bs when T != double

    for (int i = tid; i < bs * bs; i += kPcThreads) A[i] = 0.0;
    for (int i = tid; i < D * S; i += kPcThreads) sdif[i] = a.diff[D * seg + i];
    for (int i = tid; i < S; i += kPcThreads) {
        sd[i] = a.jd[seg + i];
        se[i] = (i + 1 < S) ? a.je[seg + i] : 0.0;
        spd[i] = a.pdiff[seg + i];
    }
    if (tid == 0) { misc[0] = 0; misc[1] = 0; }
    __syncthreads();

    // rows: R + B + C + P' (each thread owns one row of the block)
    const double k = a.rfrig[f];
    for (int i = tid; i < bs; i += kPcThreads) {
        const int p = i / D, d = i % D;
        if (np > 2) {
            // R = -k sum_j s_j s_j', s_j = e_j - 2 e_{j+1} + e_{j+2}
            for (int q = p - 2; q <= p + 2; ++q) {
                if (q < 0 || q >= np) continue;
                double acc = 0.0;
                for (int j = p - 2; j <= p; ++j) {
                    if (j < 0 || j > np - 3) continue;
                    const int cp = (p == j) ? 1 : (p == j + 1 ? -2 : 1);
                    const int dq = q - j;
                    if (dq < 0 || dq > 2) continue;
                    const int cq = (dq == 1) ? -2 : 1;
                    acc += (double)(cp * cq);
                }
                if (acc != 0.0) A[i + bs * (q * D + d)] += -k * acc;
            }
        }
        if (a.Buse) {
            const uint32_t r = idx + p;
            for (uint32_t q = a.Brow[r]; q < a.Brow[r + 1]; ++q) {
                const uint32_t cc = a.Bcol[q];
                if (cc >= idx && cc < idx + np) A[i + bs * ((cc - idx) * D + d)] += a.Bval[q];
            }
        }
        if (a.Cuse) {
            const uint32_t r = D * idx + i;
            for (uint32_t q = a.Crow[r]; q < a.Crow[r + 1]; ++q) {
                const uint32_t cc = a.Ccol[q];
                if (cc >= D * idx && cc < D * (idx + np)) A[i + bs * (cc - D * idx)] += a.Cval[q];
            }
        }
        if (p < S && spd[p] != 0.0) {
            A[i + bs * (p * D + d)] -= spd[p];
            A[i + bs * ((p + 1) * D + d)] += spd[p];
        }
        if (p >= 1 && spd[p - 1] != 0.0) {
            A[i + bs * (p * D + d)] -= spd[p - 1];
            A[i + bs * ((p - 1) * D + d)] += spd[p - 1];
        }
    }
    __syncthreads();

    // columns: M[:,j] = e_j - dt*mob*P A[:,j]
    const double sca = -a.dt * a.mob[f];
    for (int j = tid; j < bs; j += kPcThreads) {
        double* col = A + (size_t)bs * j;
        for (int s = 0; s < S; ++s) {
            double t = 0.0;
            for (int d = 0; d < D; ++d) t += sdif[D * s + d] * (col[D * (s + 1) + d] - col[D * s + d]);
            lam[(size_t)s * bs + j] = t;
        }
        // dptts2
        if (S == 1) {
            lam[j] *= 1.0 / sd[0];
        } else {
            double prev = lam[j];
            for (int s = 1; s < S; ++s) {
                prev = lam[(size_t)s * bs + j] - prev * se[s - 1];
                lam[(size_t)s * bs + j] = prev;
            }
            double nxt = lam[(size_t)(S - 1) * bs + j] / sd[S - 1];
            lam[(size_t)(S - 1) * bs + j] = nxt;
            for (int s = S - 2; s >= 0; --s) {
                nxt = lam[(size_t)s * bs + j] / sd[s] - nxt * se[s];
                lam[(size_t)s * bs + j] = nxt;
            }
        }
        for (int p = 0; p < np; ++p)
            for (int d = 0; d < D; ++d) {
                double v = col[D * p + d];
                if (p < S) v = v + sdif[D * p + d] * lam[(size_t)p * bs + j];
                if (p >= 1) v = v - sdif[D * (p - 1) + d] * lam[(size_t)(p - 1) * bs + j];
                col[D * p + d] = sca * v;
            }
        col[j] += 1.0;
    }
    __syncthreads();

    T* M;
    if constexpr (sizeof(T) == sizeof(double)) {
        M = (T*)A;
    } else {
        for (int i = tid; i < bs * bs; i += kPcThreads) AT[i] = (T)A[i];
        __syncthreads();
        M = AT;
    }
    if (!a.factor) {
        T* dst = out + a.blkoff[f];
        for (int i = tid; i < bs * bs; i += kPcThreads) dst[i] = M[i];
        return;
    }
    if (a.method == 0) {
        lds_gauss_jordan<T>(M, bs, ipiv, &misc[0], &misc[1], work, work + bs);
    } else {
        lds_getrf<T>(M, bs, ipiv, &misc[0], &misc[1]);
        if (misc[0] == 0) lds_getri<T>(M, bs, ipiv, work);
    }
    const int ok = (misc[0] == 0);
    T* dst = out + a.blkoff[f];
    if (ok)
        for (int i = tid; i < bs * bs; i += kPcThreads) dst[i] = M[i];
    if (tid == 0) useblk[f] = ok;
}

size_t precond_lds_bytes(int D, int bs_max, int precision)
{
    const int S = bs_max / D - 1;
    size_t b = (size_t)bs_max * bs_max * 8 + (size_t)S * bs_max * 8 + (size_t)(D + 3) * S * 8 + (size_t)bs_max * 4 +
               16 + 16 + (size_t)bs_max * 8 + 16 + 64;
    if (precision == CYMG_PRECOND_FP32) b += (size_t)bs_max * bs_max * 4;
    return b;
}

void launch_precond_build(Context& c, int precision, double* block_out)
{
    if (!c.nfib) return;
    const int D = c.dim;
    const int bs_max = D * (int)c.largest;
    const size_t lds = precond_lds_bytes(D, bs_max, precision);
    if (lds > 160 * 1024)
        throw Error(CYMG_ERR_UNSUPPORTED, "preconditioner block of " + std::to_string(bs_max) +
                                              " unknowns exceeds the LDS budget of one workgroup");
    PcArgs a;
    a.np = c.d_np.get(); a.idx = c.d_idx.get(); a.seg = c.d_seg.get(); a.blkoff = c.d_blkoff.get();
    a.mob = c.d_mob.get(); a.rfrig = c.d_rfrig.get(); a.diff = c.d_diff.get(); a.jd = c.d_jd.get();
    a.je = c.d_je.get(); a.pdiff = c.d_pdiff.get();
    a.Brow = c.mB.rowptr.get(); a.Bcol = c.mB.col.get(); a.Bval = c.mB.val.get(); a.Buse = c.mB.used;
    a.Crow = c.mC.rowptr.get(); a.Ccol = c.mC.col.get(); a.Cval = c.mC.val.get(); a.Cuse = c.mC.used;
    a.dt = c.prm.time_step;
    a.bs_max = bs_max;
    a.factor = block_out ? 0 : 1;
    a.method = c.pc_method;
    c.d_useblk.resize(c.nfib);
    const dim3 g(c.nfib), b(kPcThreads);
    if (precision == CYMG_PRECOND_FP32) {
        c.d_binv32.resize(c.blk_total);
        if (D == 3) {
            CYMG_HIP(hipFuncSetAttribute((const void*)k_precond_build<3, float>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_precond_build<3, float>), g, b, lds, c.stream, a, c.d_binv32.get(), c.d_useblk.get());
        } else {
            CYMG_HIP(hipFuncSetAttribute((const void*)k_precond_build<2, float>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_precond_build<2, float>), g, b, lds, c.stream, a, c.d_binv32.get(), c.d_useblk.get());
        }
    } else {
        double* dst = block_out;
        if (!dst) {
            c.d_binv.resize(c.blk_total);
            dst = c.d_binv.get();
        }
        if (D == 3) {
            CYMG_HIP(hipFuncSetAttribute((const void*)k_precond_build<3, double>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_precond_build<3, double>), g, b, lds, c.stream, a, dst, c.d_useblk.get());
        } else {
            CYMG_HIP(hipFuncSetAttribute((const void*)k_precond_build<2, double>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_precond_build<2, double>), g, b, lds, c.stream, a, dst, c.d_useblk.get());
        }
    }
    CYMG_HIP(hipGetLastError());
    if (!block_out) {
        c.pc_precision = precision;
        c.pc_ready = true;
    }
}

// standalone Meca::precondition
template <typename T>
__global__ __launch_bounds__(4 * kWave) void k_precond_apply(int D, const uint32_t* __restrict__ npv,
                                                             const uint32_t* __restrict__ idxv,
                                                             const uint64_t* __restrict__ blkoff,
                                                             const int* __restrict__ useblk, const T* __restrict__ B,
                                                             const double* __restrict__ X, double* __restrict__ Y)
{
    __shared__ double xs[kMaxFiberPoints * 3];
    __shared__ double red[4 * kMaxFiberPoints * 3];
    const uint32_t f = blockIdx.x;
    const int bs = D * (int)npv[f];
    const uint32_t o = D * idxv[f];
    for (int j = threadIdx.x; j < bs; j += 4 * kWave) xs[j] = X[o + j];
    __syncthreads();
    block_apply_mw<T, 4>(B + blkoff[f], bs, useblk[f] != 0, xs, Y + o, red);
}

void launch_precond_apply(Context& c, const double* X, double* Y)
{
    if (!c.nfib) return;
    const dim3 g(c.nfib), b(4 * kWave);
    if (c.pc_precision == CYMG_PRECOND_FP32)
        hipLaunchKernelGGL(k_precond_apply<float>, g, b, 0, c.stream, c.dim, c.d_np.get(), c.d_idx.get(),
                           c.d_blkoff.get(), c.d_useblk.get(), c.d_binv32.get(), X, Y);
    else
        hipLaunchKernelGGL(k_precond_apply<double>, g, b, 0, c.stream, c.dim, c.d_np.get(), c.d_idx.get(),
                           c.d_blkoff.get(), c.d_useblk.get(), c.d_binv.get(), X, Y);
    CYMG_HIP(hipGetLastError());
}

}  // namespace cymg
