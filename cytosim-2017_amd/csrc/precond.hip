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
#include <type_traits>

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
    int method;   // 0 or 2: Gauss-Jordan, 1: getrf + getri
    const uint32_t* list;   // fibers of the launch (null: all, blockIdx.x = fiber)
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
    const uint32_t f = a.list ? a.list[blockIdx.x] : blockIdx.x;
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
    if (a.method != 1) {
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

// ---------------------------------------------------------------------------
// Gauss-Jordan inverse with the block in registers: one workgroup of NW
// wavefronts per block, lane i owns row i (bs <= 64 * NW, bs <= BSMAX).  At
// step k the candidate pivots are the rows not yet used as pivot (partial
// pivoting by row search as dgetf2, meca.cc:352-353: largest magnitude, the
// magnitude compared in FP32 with its 7 low mantissa bits dropped, lowest row
// on ties within that resolution); the pivot row is scaled and broadcast
// through LDS, every other row eliminates column k.  The row array rotates
// left by one column per step, so column k is always a[0] and the inverse
// column of the step is appended at a[BSMAX-1]: all register indices are
// compile-time constants.  After bs steps, the row that was pivot of step k
// holds row k of inv(P A) with its columns in step order, i.e. row k of
// inv(A) with column j stored at perm[j] (the pivot row of step j).  A zero
// pivot marks the block unusable (meca.cc:437-438).  The assembled FP64 block
// is read from M; the inverse (type T) is written to out at the same offset.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Maximum over a wavefront with DPP row operations (quad permutations,
// half-row and row mirrors, row broadcasts 15 and 31): lane 63 ends with the
// maximum, read back as a uniform value.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));  // row_mirror
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ double readlane_t(double v, int lane)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ float readlane_t(float v, int lane)
{
    return __uint_as_float((unsigned)__builtin_amdgcn_readlane((int)__float_as_uint(v), lane));
}

template <typename T, int BSMAX, int NW>
__global__ __launch_bounds__(NW * kWave) void k_block_gj(const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ npv,
                                                         const uint64_t* __restrict__ blkoff, int D,
                                                         const double* __restrict__ M, T* __restrict__ out,
                                                         int* __restrict__ useblk)
{
    static_assert(BSMAX <= NW * kWave, "one row per lane");
    __shared__ T row[BSMAX];
    __shared__ int perm[BSMAX];
    __shared__ T wval[NW];
    __shared__ unsigned wkey[NW];
    const uint32_t f = list[blockIdx.x];
    const int bs = D * (int)npv[f];
    const uint64_t off = blkoff[f];
    const int i = threadIdx.x;
    const int lane = i & (kWave - 1);
    const int w = i / kWave;
    T a[BSMAX];
#pragma unroll
    for (int j = 0; j < BSMAX; ++j) a[j] = (j < bs && i < bs) ? (T)M[off + i + (size_t)bs * j] : T(0);
    bool used = false;
    int kst = 0, info = 0;
    for (int k = 0; k < bs; ++k) {
        const T ak = a[0];
        // candidate key: magnitude (FP32 bits, low 7 bits cleared) | 127 - row
        unsigned key = 0;
        if (i < bs && !used && ak == ak) key = (__float_as_uint((float)fabs(ak)) & ~127u) | (unsigned)(127 - i);
        unsigned best = wave_max_u32(key);
        int p;
        T piv;
        if constexpr (NW == 1) {
            p = 127 - (int)(best & 127u);
            piv = readlane_t(ak, best ? p : 0);
        } else {
            const int pw = 127 - (int)(best & 127u);
            if (lane == 0) wkey[w] = best;
            if (best && i == pw) wval[w] = ak;
            __syncthreads();
            int ww = 0;
            best = wkey[0];
#pragma unroll
            for (int q = 1; q < NW; ++q)
                if (wkey[q] > best) { best = wkey[q]; ww = q; }
            p = 127 - (int)(best & 127u);
            piv = best ? wval[ww] : T(0);
        }
        if (best == 0u || piv == T(0)) {
            info = k + 1;
            break;
        }
        const T r = T(1) / piv;
        const bool isp = (i == p);
        if (isp) {
            // the scaled pivot row goes to LDS; the registers are replaced by it
            // in the update below (a divergent in-place scaling would keep two
            // copies of the row live)
#pragma unroll
            for (int j = 1; j < BSMAX; ++j) row[j - 1] = a[j] * r;
            used = true;
            kst = k;
        }
        if (i == 0) perm[k] = p;
        if constexpr (NW == 1) wave_sync();
        else __syncthreads();
        const T m = ak;
#pragma unroll
        for (int j = 1; j < BSMAX; ++j) {
            const T rj = row[j - 1];
            a[j - 1] = isp ? rj : a[j] - m * rj;
        }
        a[BSMAX - 1] = isp ? r : -m * r;
    }
    if constexpr (NW == 1) wave_sync();
    else __syncthreads();
    if (info == 0 && i < bs) {
        T* dst = out + off;
#pragma unroll
        for (int t = 0; t < BSMAX; ++t) {
            const int j = t - (BSMAX - bs);
            if (j >= 0) dst[kst + (size_t)bs * perm[j]] = a[t];
        }
    }
    if (i == 0) useblk[f] = (info == 0);
}

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inverse on the matrix cores.  One workgroup of 4
// wavefronts per block; the block lives in LDS, column-major with leading
// dimension ld (odd), padded to nb = 16*ceil(bs/16) <= 96 with an identity
// tail.  For each panel K of 16 columns:
//  * wavefront 0 runs the panel's elimination steps on the panel columns only,
//    rows in registers (lanes l and l + 64).  Pivot choice as k_block_gj
//    (largest FP32-truncated magnitude among the rows not yet pivot, lowest
//    row on ties; no interchanges, the pivot row of each step is recorded);
//  * the other columns J receive the panel's steps at once as a product on
//    the matrix cores (v_mfma_{f64,f32}_16x16x4, 16x16 output tiles, K = 16 in
//    4 steps):  A[:, J] <- Z A[:, J] + X A[P, J],  where X = A[:, K] after the
//    panel steps, P the panel's pivot rows in step order and Z the identity
//    with the rows P zeroed (blocked Gauss-Jordan inversion: Quintana-Orti,
//    Quintana-Orti, Sun and van de Geijn, SIAM J. Sci. Comput. 23 (2001)).
// After the last panel the row that was pivot of step s holds row s of
// inv(A), its column t at the pivot row of step t; the result is written in
// natural order, coalesced.  A zero pivot marks the block unusable
// (meca.cc:437-438).  Reads the assembled FP64 block from M; writes the
// inverse (type T, factorised in T) to out at the same offset.
constexpr int kInvNB = 96;
constexpr int kInvThreads = 256;

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
// Output layout of the 16x16x4 products (measured on gfx950): lane l holds
// column l % 16; its register r holds row 4 (l / 16) + r for FP32 and row
// l / 16 + 4 r for FP64.
template <typename T> struct MfmaVec;
template <> struct MfmaVec<double> {
    typedef f64x4 type;
    static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <> struct MfmaVec<float> {
    typedef f32x4v type;
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

__device__ __forceinline__ f64x4 mfma_16x16x4(double a, double b, f64x4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4v mfma_16x16x4(float a, float b, f32x4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// The elimination steps of one panel (columns k .. k + steps - 1) in
// wavefront 0.  TWO: rows 64..nb-1 exist (held in v1).  Returns false on a
// zero pivot (sets *info).
template <typename T, bool TWO>
__device__ __forceinline__ bool inv_panel(T* A, int ld, int nb, int bs, int k, int steps, unsigned& used, int* perm,
                                          int* stepof, int* info)
{
    const int lane = threadIdx.x & (kWave - 1);
    const bool row0 = lane < nb, row1 = TWO && lane + 64 < nb;   // the lane's rows inside the block
    T v0[16], v1[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        v0[c] = row0 ? A[lane + ld * (k + c)] : T(0);
        v1[c] = row1 ? A[lane + 64 + ld * (k + c)] : T(0);
    }
    // fully unrolled (every register index a constant): no early exit, the
    // steps past `steps` or after a zero pivot are skipped by a uniform test
    bool ok = true;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        if (!ok || t >= steps) continue;
        const T a0 = v0[t], a1 = v1[t];
        unsigned k0 = 0, k1 = 0;
        if (lane < bs && !(used & 1u) && a0 == a0)
            k0 = (__float_as_uint((float)fabs(a0)) & ~127u) | (unsigned)(127 - lane);
        if (TWO && lane + 64 < bs && !(used & 2u) && a1 == a1)
            k1 = (__float_as_uint((float)fabs(a1)) & ~127u) | (unsigned)(127 - (lane + 64));
        const unsigned best = wave_max_u32(max(k0, k1));
        const int p = 127 - (int)(best & 127u);
        const int pl = p & (kWave - 1);
        const bool hi = TWO && p >= 64;
        const T piv = readlane_t(hi ? v1[t] : v0[t], pl);
        if (best == 0u || piv == T(0)) {
            if (lane == 0) *info = k + t + 1;
            ok = false;
            continue;
        }
        const T r = T(1) / piv;
        const T m0 = v0[t], m1 = v1[t];
        const bool own0 = lane == pl && !hi, own1 = lane == pl && hi;
        // column c of the scaled pivot row is read from lane pl before the
        // lanes update column c; the pivot lane takes it in place of its row
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            if (c == t) continue;
            const T sc = readlane_t(hi ? v1[c] : v0[c], pl) * r;
            const T n0 = v0[c] - m0 * sc;
            v0[c] = own0 ? sc : n0;
            if (TWO) {
                const T n1 = v1[c] - m1 * sc;
                v1[c] = own1 ? sc : n1;
            }
        }
        v0[t] = own0 ? r : -m0 * r;
        if (TWO) v1[t] = own1 ? r : -m1 * r;
        if (own0) used |= 1u;
        if (own1) used |= 2u;
        if (lane == 0) {
            perm[k + t] = p;
            stepof[p] = k + t;
        }
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        if (row0) A[lane + ld * (k + c)] = v0[c];
        if (row1) A[lane + 64 + ld * (k + c)] = v1[c];
    }
    return ok;
}

template <typename T>
__global__ __launch_bounds__(kInvThreads) void k_block_inv_mfma(const uint32_t* __restrict__ list,
                                                                const uint32_t* __restrict__ npv,
                                                                const uint64_t* __restrict__ blkoff, int D, int ld,
                                                                const double* __restrict__ M, T* __restrict__ out,
                                                                int* __restrict__ useblk)
{
    typedef typename MfmaVec<T>::type V4;
    extern __shared__ double inv_lds[];
    T* A = (T*)inv_lds;                                   // ld * (ld - 1)
    int* perm = (int*)(A + (size_t)ld * (ld - 1));        // pivot row of each step
    int* stepof = perm + kInvNB;                          // step at which each row was pivot
    int* info = stepof + kInvNB;
    const uint32_t f = list[blockIdx.x];
    const int bs = D * (int)npv[f];
    const int nb = (bs + 15) & ~15;
    const uint64_t off = blkoff[f];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;

    for (int j = w; j < nb; j += kInvThreads / kWave)
        for (int i = lane; i < nb; i += kWave)
            A[i + ld * j] = (i < bs && j < bs) ? (T)M[off + i + (size_t)bs * j] : (i == j ? T(1) : T(0));
    if (tid < kInvNB) { perm[tid] = tid; stepof[tid] = -1; }
    if (tid == 0) *info = 0;
    __syncthreads();

    const int nt = nb / 16;
    unsigned used = 0;
    for (int k = 0; k < bs; k += 16) {
        const int steps = min(16, bs - k);
        if (w == 0) {
            if (nb > 64) inv_panel<T, true>(A, ld, nb, bs, k, steps, used, perm, stepof, info);
            else inv_panel<T, false>(A, ld, nb, bs, k, steps, used, perm, stepof, info);
        }
        __syncthreads();
        if (*info) break;
        const int kt = k / 16;
        // B operands: the panel's pivot rows (step order) in every other column tile
        T bop[kInvNB / 16][4];
        const int s0 = lane >> 4, jl = lane & 15;
#pragma unroll
        for (int jt = 0; jt < kInvNB / 16; ++jt)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int s = 4 * c + s0;
                bop[jt][c] = (jt < nt && jt != kt && s < steps) ? A[perm[k + s] + ld * (16 * jt + jl)] : T(0);
            }
        __syncthreads();
        int q = 0;
#pragma unroll
        for (int jt = 0; jt < kInvNB / 16; ++jt) {
            if (jt >= nt || jt == kt) continue;
            for (int it = 0; it < nt; ++it, ++q) {
                if ((q & 3) != w) continue;
                const int j = 16 * jt + jl;
                V4 acc;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * it + MfmaVec<T>::row(s0, r);
                    const int st = stepof[i];
                    acc[r] = (st >= k && st < k + steps) ? T(0) : A[i + ld * j];
                }
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    acc = mfma_16x16x4(A[16 * it + jl + ld * (k + 4 * c + s0)], bop[jt][c], acc);
#pragma unroll
                for (int r = 0; r < 4; ++r) A[16 * it + MfmaVec<T>::row(s0, r) + ld * j] = acc[r];
            }
        }
        __syncthreads();
    }
    const bool ok = (*info == 0);
    if (ok) {
        T* dst = out + off;
        for (int cidx = w; cidx < bs; cidx += kInvThreads / kWave) {
            const int t = stepof[cidx];
            for (int s = lane; s < bs; s += kWave) dst[s + (size_t)bs * cidx] = A[perm[s] + ld * t];
        }
    }
    if (tid == 0) useblk[f] = ok;
}

static size_t inv_mfma_lds_bytes(int nbmax, size_t elem)
{
    const int ld = nbmax + 1;
    return (size_t)ld * nbmax * elem + (2 * kInvNB + 4) * sizeof(int);
}

// LDS of one workgroup on gfx950
constexpr size_t kLdsBudget = 160 * 1024;
// columns per panel of the blocked Gauss-Jordan of the blocks beyond LDS
constexpr int kBigPanel = 16;

// a block the workgroup-local kernels take (assembly and inversion in LDS, in
// either precision); larger blocks are assembled column by column and
// inverted in device memory
static bool lds_block(int D, int bs)
{
    return precond_lds_bytes(D, bs, CYMG_PRECOND_FP32) <= kLdsBudget;
}

// Fibers grouped by block size: [<=32 | <=64 | <=96 | <=128 | larger in LDS |
// beyond LDS], made once per layout.  Stored as [<=96 | <=64 | <=32 | <=128 |
// larger | beyond] so that the matrix-core kernel takes the blocks up to 96
// unknowns in one range, largest class first, and the classes the LDS
// kernels take are one range from the start.
static void gj_lists(Context& c)
{
    if (c.gj_gen == c.layout_gen && !c.gj_start.empty()) return;
    const int edges[4] = { 32, 64, 96, 128 };
    std::vector<uint32_t> lists[6];
    int bs_lds = 0;
    for (uint32_t f = 0; f < c.nfib; ++f) {
        const int bs = c.dim * (int)c.h_np[f];
        int b = 0;
        if (!lds_block(c.dim, bs)) {
            b = 5;
        } else {
            while (b < 4 && bs > edges[b]) ++b;
            bs_lds = std::max(bs_lds, bs);
        }
        lists[b].push_back(f);
    }
    std::vector<uint32_t> all;
    c.gj_start.assign(12, 0);   // [b] start, [6 + b] end of class b
    const int order[6] = { 2, 1, 0, 3, 4, 5 };
    for (int b : order) {
        c.gj_start[b] = (uint32_t)all.size();
        all.insert(all.end(), lists[b].begin(), lists[b].end());
        c.gj_start[6 + b] = (uint32_t)all.size();
    }
    c.d_gj_list.resize(all.size() ? all.size() : 1);
    if (!all.empty()) {
        CYMG_HIP(hipMemcpyAsync(c.d_gj_list.get(), all.data(), all.size() * 4, hipMemcpyHostToDevice, c.stream));
        CYMG_HIP(hipStreamSynchronize(c.stream));   // pageable source
    }
    c.gj_gen = c.layout_gen;
    c.gj_bs_lds = bs_lds;
    // work space of the blocks beyond LDS, by position in their class
    std::vector<uint64_t> sofs;
    std::vector<uint32_t> iofs;
    uint64_t sa = 0;
    uint32_t si = 0, bmax = 0;
    for (uint32_t f : lists[5]) {
        const uint32_t bs = c.dim * c.h_np[f];
        sofs.push_back(sa);
        iofs.push_back(si);
        sa += (uint64_t)bs * bs;
        si += bs;
        bmax = std::max(bmax, bs);
    }
    c.big_bsmax = bmax;
    if (!sofs.empty()) {
        c.d_big_sofs.resize(sofs.size());
        c.d_big_iofs.resize(iofs.size());
        c.d_big_A.resize(sa);
        c.d_big_AP.resize((size_t)kBigPanel * si);
        c.d_big_piv.resize(si);
        c.d_big_step.resize(si);
        c.d_big_fail.resize(sofs.size());
        CYMG_HIP(hipMemcpyAsync(c.d_big_sofs.get(), sofs.data(), sofs.size() * 8, hipMemcpyHostToDevice, c.stream));
        CYMG_HIP(hipMemcpyAsync(c.d_big_iofs.get(), iofs.data(), iofs.size() * 4, hipMemcpyHostToDevice, c.stream));
        CYMG_HIP(hipStreamSynchronize(c.stream));   // pageable sources
    }
}

// Gauss-Jordan inverse of the blocks beyond LDS, blocked by panels of
// kBigPanel columns, in device memory.  Pivot rows are chosen among the rows
// not yet used (largest magnitude in the column, the first on ties), without
// row interchanges: after all steps the array holds the inverse with its rows
// and columns permuted, inverse(s, c) = A(piv[s], step[c]).  A panel is
// eliminated by one workgroup (k_big_panel); its transformation reaches every
// other column as one product over the panel, A(:, j) <- [i not pivot of the
// panel] A(:, j) + W A(P, j) with W the eliminated panel and P its pivot rows,
// spread over many workgroups (k_big_update).  A zero or NaN pivot leaves the
// block unused (identity), as the workgroup-local kernels do.
constexpr int kBigThreads = 512;
constexpr int kBigMaxBs = CYMG_MAX_FIBER_POINTS * 3;
constexpr int kBigRows = 256, kBigCols = 64;

struct BigArgs {
    const uint32_t* list;
    const uint32_t* np;
    const uint64_t* sofs;
    const uint32_t* iofs;
    double* A;
    double* AP;
    int* piv;
    int* step;
    int* fail;
    int D;
};

__global__ __launch_bounds__(256) void k_big_init(BigArgs g)
{
    const uint32_t b = blockIdx.x;
    const int bs = g.D * (int)g.np[g.list[b]];
    int* step = g.step + g.iofs[b];
    for (int i = threadIdx.x; i < bs; i += blockDim.x) step[i] = bs;   // bs: not yet a pivot row
    if (threadIdx.x == 0) g.fail[b] = 0;
}

__global__ __launch_bounds__(kBigThreads) void k_big_panel(BigArgs g, int k0)
{
    __shared__ double col[kBigMaxBs];
    __shared__ double rv[kBigThreads / kWave];
    __shared__ int ri[kBigThreads / kWave];
    __shared__ int sp[kBigPanel];
    __shared__ double apj[kBigPanel];
    constexpr int NW = kBigThreads / kWave;
    const uint32_t b = blockIdx.x;
    const int bs = g.D * (int)g.np[g.list[b]];
    if (k0 >= bs || g.fail[b]) return;
    double* A = g.A + g.sofs[b];
    int* step = g.step + g.iofs[b];
    int* piv = g.piv + g.iofs[b];
    const int m = bs - k0 < kBigPanel ? bs - k0 : kBigPanel;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    for (int q = 0; q < m; ++q) {
        const int k = k0 + q;
        double bv = -1.0;
        int bi = bs;
        for (int i = tid; i < bs; i += kBigThreads) {
            if (step[i] != bs) continue;
            const double v = fabs(A[i + (size_t)bs * k]);
            if (v > bv || (v != v && bv == bv)) { bv = v; bi = i; }
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const double ov = __shfl_xor(bv, o, kWave);
            const int oi = __shfl_xor(bi, o, kWave);
            if (ov > bv || (ov == bv && oi < bi) || (ov != ov && bv == bv)) { bv = ov; bi = oi; }
        }
        if (lane == 0) { rv[w] = bv; ri[w] = bi; }
        __syncthreads();
        bv = rv[0];
        bi = ri[0];
        for (int t = 1; t < NW; ++t)
            if (rv[t] > bv || (rv[t] == bv && ri[t] < bi) || (rv[t] != rv[t] && bv == bv)) { bv = rv[t]; bi = ri[t]; }
        if (!(bv > 0.0)) {
            if (tid == 0) g.fail[b] = k + 1;
            return;
        }
        const int p = bi;
        for (int i = tid; i < bs; i += kBigThreads) col[i] = A[i + (size_t)bs * k];
        if (tid < m) apj[tid] = A[p + (size_t)bs * (k0 + tid)];
        __syncthreads();
        const double r = 1.0 / col[p];
        for (int i = tid; i < bs; i += kBigThreads) {
            const double ci = col[i];
            for (int t = 0; t < m; ++t) {
                double* cj = A + (size_t)bs * (k0 + t);
                if (t == q) cj[i] = (i == p) ? r : -ci * r;
                else {
                    const double a = apj[t] * r;
                    cj[i] = (i == p) ? a : cj[i] - ci * a;
                }
            }
        }
        if (tid == 0) { step[p] = k; piv[k] = p; sp[q] = p; }
        __syncthreads();
    }
    // the panel's pivot rows of every column, before the update overwrites them
    double* AP = g.AP + (size_t)kBigPanel * g.iofs[b];
    for (int t = 0; t < m; ++t)
        for (int j = tid; j < bs; j += kBigThreads) AP[(size_t)bs * t + j] = A[sp[t] + (size_t)bs * j];
}

__global__ __launch_bounds__(kBigRows) void k_big_update(BigArgs g, int k0)
{
    __shared__ double aps[kBigPanel][kBigCols];
    const uint32_t b = blockIdx.x;
    const int bs = g.D * (int)g.np[g.list[b]];
    const int j0 = blockIdx.y * kBigCols, i = blockIdx.z * kBigRows + threadIdx.x;
    if (k0 >= bs || j0 >= bs || blockIdx.z * kBigRows >= (unsigned)bs || g.fail[b]) return;
    double* A = g.A + g.sofs[b];
    const int* step = g.step + g.iofs[b];
    const double* AP = g.AP + (size_t)kBigPanel * g.iofs[b];
    const int m = bs - k0 < kBigPanel ? bs - k0 : kBigPanel;
    for (int e = threadIdx.x; e < kBigPanel * kBigCols; e += kBigRows) {
        const int t = e / kBigCols, jj = e % kBigCols;
        aps[t][jj] = (t < m && j0 + jj < bs) ? AP[(size_t)bs * t + j0 + jj] : 0.0;
    }
    __syncthreads();
    if (i >= bs) return;
    double W[kBigPanel];
#pragma unroll
    for (int t = 0; t < kBigPanel; ++t) W[t] = t < m ? A[i + (size_t)bs * (k0 + t)] : 0.0;
    const bool pivot_row = step[i] >= k0 && step[i] < k0 + m;
    const int je = bs - j0 < kBigCols ? bs - j0 : kBigCols;
    for (int jj = 0; jj < je; ++jj) {
        const int j = j0 + jj;
        if (j >= k0 && j < k0 + m) continue;
        double v = pivot_row ? 0.0 : A[i + (size_t)bs * j];
#pragma unroll
        for (int t = 0; t < kBigPanel; ++t) v += W[t] * aps[t][jj];
        A[i + (size_t)bs * j] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_big_store(BigArgs g, const uint64_t* __restrict__ blkoff,
                                                   T* __restrict__ out, int* __restrict__ useblk)
{
    const uint32_t b = blockIdx.x, f = g.list[b];
    const int bs = g.D * (int)g.np[f];
    const bool ok = g.fail[b] == 0;
    if (blockIdx.y == 0 && threadIdx.x == 0) useblk[f] = ok;
    if (!ok) return;
    const double* A = g.A + g.sofs[b];
    const int* step = g.step + g.iofs[b];
    const int* piv = g.piv + g.iofs[b];
    T* dst = out + blkoff[f];
    for (int cc = blockIdx.y; cc < bs; cc += gridDim.y) {
        const double* src = A + (size_t)bs * step[cc];
        for (int s2 = threadIdx.x; s2 < bs; s2 += blockDim.x) dst[s2 + (size_t)bs * cc] = (T)src[piv[s2]];
    }
}

// blocks list[begin .. begin + count) of the class beyond LDS: assembled
// column by column (into M by fiber offsets when only the blocks are asked
// for), or assembled into the work space and inverted into d_binv / d_binv32
static void launch_big(Context& c, uint32_t begin, uint32_t count, int precision, bool invert, double* M)
{
    if (!count) return;
    const uint32_t* list = c.d_gj_list.get() + begin;
    if (!invert) {
        launch_block_columns(c, list, count, c.d_blkoff.get(), 0, M);
        return;
    }
    launch_block_columns(c, list, count, c.d_big_sofs.get(), 1, c.d_big_A.get());
    BigArgs g;
    g.list = list; g.np = c.d_np.get(); g.sofs = c.d_big_sofs.get(); g.iofs = c.d_big_iofs.get();
    g.A = c.d_big_A.get(); g.AP = c.d_big_AP.get(); g.piv = c.d_big_piv.get(); g.step = c.d_big_step.get();
    g.fail = c.d_big_fail.get(); g.D = c.dim;
    const int bsmax = (int)c.big_bsmax;
    hipLaunchKernelGGL(k_big_init, dim3(count), dim3(256), 0, c.stream, g);
    const dim3 ug(count, (bsmax + kBigCols - 1) / kBigCols, (bsmax + kBigRows - 1) / kBigRows);
    for (int k0 = 0; k0 < bsmax; k0 += kBigPanel) {
        hipLaunchKernelGGL(k_big_panel, dim3(count), dim3(kBigThreads), 0, c.stream, g, k0);
        hipLaunchKernelGGL(k_big_update, ug, dim3(kBigRows), 0, c.stream, g, k0);
    }
    const dim3 sg(count, std::min(bsmax, 64));
    if (precision == CYMG_PRECOND_FP32)
        hipLaunchKernelGGL(k_big_store<float>, sg, dim3(256), 0, c.stream, g, c.d_blkoff.get(), c.d_binv32.get(),
                           c.d_useblk.get());
    else
        hipLaunchKernelGGL(k_big_store<double>, sg, dim3(256), 0, c.stream, g, c.d_blkoff.get(), c.d_binv.get(),
                           c.d_useblk.get());
    CYMG_HIP(hipGetLastError());
}

template <typename T, int BSMAX, int NW>
static void launch_gj(Context& c, uint32_t begin, uint32_t end, T* out)
{
    if (end <= begin) return;
    hipLaunchKernelGGL((k_block_gj<T, BSMAX, NW>), dim3(end - begin), dim3(NW * kWave), 0, c.stream,
                       c.d_gj_list.get() + begin, c.d_np.get(), c.d_blkoff.get(), c.dim, c.d_binv.get(), out,
                       c.d_useblk.get());
}

// blocks list[begin .. begin + count) (at most 96 unknowns, largest first)
template <typename T>
static void launch_inv_mfma(Context& c, uint32_t begin, uint32_t count, T* out)
{
    if (!count) return;
    int nbmax = 0;
    for (uint32_t f = 0; f < c.nfib; ++f) {
        const int bs = c.dim * (int)c.h_np[f];
        if (bs <= kInvNB) nbmax = std::max(nbmax, (bs + 15) & ~15);
    }
    const size_t lds = inv_mfma_lds_bytes(nbmax, sizeof(T));
    CYMG_HIP(hipFuncSetAttribute((const void*)k_block_inv_mfma<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
    hipLaunchKernelGGL(k_block_inv_mfma<T>, dim3(count), dim3(kInvThreads), lds, c.stream,
                       c.d_gj_list.get() + begin, c.d_np.get(), c.d_blkoff.get(), c.dim, nbmax + 1,
                       c.d_binv.get(), out, c.d_useblk.get());
}

static void launch_lds_build(Context& c, const PcArgs& a0, int precision, void* dst, const uint32_t* list,
                             uint32_t count)
{
    if (!count) return;
    PcArgs a = a0;
    a.list = list;
    const int D = c.dim;
    const size_t lds = precond_lds_bytes(D, a.bs_max, precision);
    const dim3 g(count), b(kPcThreads);
    if (precision == CYMG_PRECOND_FP32) {
        auto kern = D == 3 ? k_precond_build<3, float> : k_precond_build<2, float>;
        CYMG_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, g, b, lds, c.stream, a, (float*)dst, c.d_useblk.get());
    } else {
        auto kern = D == 3 ? k_precond_build<3, double> : k_precond_build<2, double>;
        CYMG_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, g, b, lds, c.stream, a, (double*)dst, c.d_useblk.get());
    }
}

void launch_precond_build(Context& c, int precision, double* block_out)
{
    if (!block_out) launch_solid_pc_build(c);
    if (!c.nfib) {
        c.pc_precision = precision;
        c.pc_ready = true;
        return;
    }
    gj_lists(c);
    const std::vector<uint32_t>& st = c.gj_start;   // [b] start, [6 + b] end of class b
    const uint32_t nlds = st[6 + 4];                 // classes 0-4: one range from the start
    const uint32_t big0 = st[5], nbig = st[6 + 5] - st[5];
    PcArgs a;
    a.np = c.d_np.get(); a.idx = c.d_idx.get(); a.seg = c.d_seg.get(); a.blkoff = c.d_blkoff.get();
    a.mob = c.d_mob.get(); a.rfrig = c.d_rfrig.get(); a.diff = c.d_diff.get(); a.jd = c.d_jd.get();
    a.je = c.d_je.get(); a.pdiff = c.d_pdiff.get();
    a.Brow = c.mB.rowptr.get(); a.Bcol = c.mB.col.get(); a.Bval = c.mB.val.get(); a.Buse = c.mB.used;
    a.Crow = c.mC.rowptr.get(); a.Ccol = c.mC.col.get(); a.Cval = c.mC.val.get(); a.Cuse = c.mC.used;
    a.dt = c.prm.time_step;
    a.bs_max = c.gj_bs_lds;
    a.factor = 1;
    a.method = c.pc_method;
    a.list = nullptr;
    c.d_useblk.resize(c.nfib);
    const uint32_t* lds_list = c.d_gj_list.get();
    if (block_out) {
        // the assembled blocks themselves
        a.factor = 0;
        launch_lds_build(c, a, CYMG_PRECOND_FP64, block_out, lds_list, nlds);
        launch_big(c, big0, nbig, CYMG_PRECOND_FP64, false, block_out);
        CYMG_HIP(hipGetLastError());
        return;
    }
    c.d_binv.resize(c.blk_total);
    if (precision == CYMG_PRECOND_FP32) c.d_binv32.resize(c.blk_total);
    void* dst = precision == CYMG_PRECOND_FP32 ? (void*)c.d_binv32.get() : (void*)c.d_binv.get();
    if (c.pc_method == 1 || c.pc_method == 2) {
        // LDS kernels for every block they take: 1 getrf + getri, 2 Gauss-Jordan
        launch_lds_build(c, a, precision, dst, lds_list, nlds);
    } else {
        // assemble the blocks into d_binv, invert by size class (method 0:
        // in registers, method 3: on the matrix cores); the classes beyond
        // are rebuilt and inverted in LDS
        PcArgs e = a;
        e.factor = 0;
        launch_lds_build(c, e, CYMG_PRECOND_FP64, c.d_binv.get(), lds_list, nlds);
        if (c.pc_method == 3) {
            const uint32_t n96 = st[6 + 0] - st[2];
            if (precision == CYMG_PRECOND_FP32) launch_inv_mfma<float>(c, st[2], n96, c.d_binv32.get());
            else launch_inv_mfma<double>(c, st[2], n96, c.d_binv.get());
            launch_lds_build(c, a, precision, dst, c.d_gj_list.get() + st[3], st[6 + 4] - st[3]);
        } else if (precision == CYMG_PRECOND_FP32) {
            float* o = c.d_binv32.get();
            launch_gj<float, 32, 1>(c, st[0], st[6 + 0], o);
            launch_gj<float, 64, 1>(c, st[1], st[6 + 1], o);
            launch_gj<float, 96, 2>(c, st[2], st[6 + 2], o);
            launch_gj<float, 128, 2>(c, st[3], st[6 + 3], o);
            launch_lds_build(c, a, precision, dst, c.d_gj_list.get() + st[4], st[6 + 4] - st[4]);
        } else {
            double* o = c.d_binv.get();
            launch_gj<double, 32, 1>(c, st[0], st[6 + 0], o);
            launch_gj<double, 64, 1>(c, st[1], st[6 + 1], o);
            launch_gj<double, 96, 2>(c, st[2], st[6 + 2], o);
            launch_lds_build(c, a, precision, dst, c.d_gj_list.get() + st[3], st[6 + 4] - st[3]);
        }
    }
    launch_big(c, big0, nbig, precision, true, c.d_binv.get());
    CYMG_HIP(hipGetLastError());
    c.pc_precision = precision;
    c.pc_ready = true;
}

// standalone Meca::precondition
template <typename T>
__global__ __launch_bounds__(4 * kWave) void k_precond_apply(int D, const uint32_t* __restrict__ npv,
                                                             const uint32_t* __restrict__ idxv,
                                                             const uint64_t* __restrict__ blkoff,
                                                             const int* __restrict__ useblk, const T* __restrict__ B,
                                                             const double* __restrict__ X, double* __restrict__ Y,
                                                             int f32apply)
{
    __shared__ double xs[kMaxFiberPoints * 3];
    __shared__ double red[4 * kMaxFiberPoints * 3];
    const uint32_t f = blockIdx.x;
    const int bs = D * (int)npv[f];
    const uint32_t o = D * idxv[f];
    for (int j = threadIdx.x; j < bs; j += 4 * kWave) xs[j] = X[o + j];
    __syncthreads();
    if constexpr (std::is_same<T, float>::value) {
        if (f32apply) {
            block_apply_mw_f32<4>(B + blkoff[f], bs, useblk[f] != 0, xs, Y + o, red);
            return;
        }
    }
    block_apply_mw<T, 4>(B + blkoff[f], bs, useblk[f] != 0, xs, Y + o, red);
}

void launch_precond_apply(Context& c, const double* X, double* Y)
{
    launch_solid_pc_apply(c, X, Y);
    if (!c.nfib) return;
    const dim3 g(c.nfib), b(4 * kWave);
    if (c.pc_precision == CYMG_PRECOND_FP32)
        hipLaunchKernelGGL(k_precond_apply<float>, g, b, 0, c.stream, c.dim, c.d_np.get(), c.d_idx.get(),
                           c.d_blkoff.get(), c.d_useblk.get(), c.d_binv32.get(), X, Y, c.fp32_apply);
    else
        hipLaunchKernelGGL(k_precond_apply<double>, g, b, 0, c.stream, c.dim, c.d_np.get(), c.d_idx.get(),
                           c.d_blkoff.get(), c.d_useblk.get(), c.d_binv.get(), X, Y, 0);
    CYMG_HIP(hipGetLastError());
}

}  // namespace cymg
