// Device-side operators over a tile: a run of consecutive whole fibers of one
// system (at most kTileMaxPoints points and kTileMaxFibers fibers) handled by
// one workgroup of kTileThreads threads.  Thread u owns the tile's unknowns u,
// u + kTileThreads, ...  The tile's slice of the operand, its forces and its
// result stay in LDS; the coupling rows of mB/mC are gathered from HBM.
//
// Grouping fibers this way keeps every lane busy on short fibers (a 20-point
// fiber has 60 unknowns, a quarter of a 256-thread workgroup) and divides the
// number of per-workgroup partial sums the Krylov launches re-reduce.
#pragma once

#include "fiber_ops.h"

namespace cymg {

constexpr int kTileThreads = 256;
constexpr int kTileWaves = kTileThreads / kWave;
constexpr int kTileMaxPoints = 170;   // 510 unknowns in 3D: one per thread of the 512-thread launches
constexpr int kTileMaxFibers = 64;

// The tiling (TileDesc, one per tile, read with one scalar load).
struct TileArgs {
    const TileDesc* desc;
    const uint32_t* fslot;   // per fiber: first lane of its segments in the packed layout
    uint32_t ntiles;
};

template <int D>
struct TileLds {
    double x[kTileMaxPoints * D];      // operand slice
    double f[kTileMaxPoints * D];      // forces, then free for the caller
    double y[kTileMaxPoints * D];      // operator result
    double lam[kTileMaxPoints];        // Lagrange multipliers, one per segment
    double pa[kTileMaxPoints], pb[kTileMaxPoints], pc[kTileMaxPoints];
    double fk[kTileMaxFibers];         // rfRigidity of each fiber of the tile
    double fsca[kTileMaxFibers];       // -dt * rfMobility
    uint64_t fblk[kTileMaxFibers];     // offset of the preconditioner block
    uint32_t fo[kTileMaxFibers + 1];   // first point of each fiber, tile-local; fo[nft] = points
    uint32_t fseg[kTileMaxFibers];     // first segment of each fiber (global)
    int fuse[kTileMaxFibers];          // block usable
    int maxS;                          // most segments of one fiber of the tile
};

struct Tile {
    uint32_t t, fb, nft, P0, tp, ts, U0, nu, sys, tb, te, S0, ws;
};

template <int D>
__device__ __forceinline__ Tile tile_begin(const TileArgs& ta, uint32_t t)
{
    const TileDesc d = ta.desc[t];
    Tile T;
    T.t = t;
    T.fb = d.fb;
    T.nft = d.nft;
    T.P0 = d.P0;
    T.tp = d.tp;
    T.ts = d.tp - d.nft;
    T.U0 = D * d.P0;
    T.nu = D * d.tp;
    T.sys = d.sys;
    T.tb = d.tb;
    T.te = d.te;
    T.S0 = d.S0;
    T.ws = d.ws;
    return T;
}

// Per-fiber tables of the tile (wave 0; nft <= 64).  The caller synchronises.
template <bool BLK, typename LDS>
__device__ __forceinline__ void tile_tables(const FiberArgs& a, LDS& L, const Tile& T,
                                            const uint64_t* blkoff, const int* useblk)
{
    const int k = threadIdx.x;
    if (k >= kWave) return;
    int S = 0;
    if (k < (int)T.nft) {
        const uint32_t f = T.fb + k;
        L.fo[k] = a.idx[f] - T.P0;
        L.fseg[k] = a.seg[f];
        L.fk[k] = a.rfrig[f];
        L.fsca[k] = -a.dt * a.mob[f];
        if (BLK) {
            L.fblk[k] = blkoff[f];
            L.fuse[k] = useblk[f];
        }
        S = (int)a.np[f] - 1;
    }
    if (k == 0) L.fo[T.nft] = T.tp;
    S = wave_max_i32(S);
    if (k == 0) L.maxS = S;
}

// tile fiber holding tile point pt: the last k with fo[k] <= pt
__device__ __forceinline__ int tile_fiber_of(const uint32_t* fo, int nft, uint32_t pt)
{
    int lo = 0, hi = nft;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (fo[mid] <= pt) lo = mid;
        else hi = mid;
    }
    return lo;
}

// tile fiber holding tile segment s (fiber k's segments start at fo[k] - k)
__device__ __forceinline__ int tile_fiber_of_seg(const uint32_t* fo, int nft, uint32_t s)
{
    int lo = 0, hi = nft;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (fo[mid] - (uint32_t)mid <= s) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The rows a thread owns: tile fiber and point within the fiber of each.
template <int D>
struct TileRows {
    static constexpr int R = (kTileMaxPoints * D + kTileThreads - 1) / kTileThreads;
    int k[R];    // -1: no row
    int p[R];
};

template <int D>
__device__ __forceinline__ void tile_rows(const TileLds<D>& L, const Tile& T, TileRows<D>& rw)
{
#pragma unroll
    for (int q = 0; q < TileRows<D>::R; ++q) {
        const int u = threadIdx.x + q * kTileThreads;
        rw.k[q] = -1;
        rw.p[q] = 0;
        if (u < (int)T.nu) {
            const uint32_t pt = (uint32_t)u / D;
            const int k = tile_fiber_of(L.fo, (int)T.nft, pt);
            rw.k[q] = k;
            rw.p[q] = (int)(pt - L.fo[k]);
        }
    }
}

// Parallel cyclic reduction of the block-diagonal tridiagonal system of the
// tile's segments in LDS (one thread per segment, ts <= kTileThreads); the
// couplings across fiber ends are zero, so maxS bounds the number of steps.
// The same recurrence as pcr_solve, with identity rows outside [0, ts).
__device__ __forceinline__ double pcr_lds(double* pa, double* pb, double* pc, double* pd, int ts, int maxS,
                                          double a, double b, double c, double d)
{
    const int s = threadIdx.x;
    for (int st = 1; st < maxS; st <<= 1) {
        if (s < ts) { pa[s] = a; pb[s] = b; pc[s] = c; pd[s] = d; }
        __syncthreads();
        if (s < ts) {
            double al = 0.0, bl = 1.0, cl = 0.0, dl = 0.0, au = 0.0, bu = 1.0, cu = 0.0, du = 0.0;
            if (s >= st) { al = pa[s - st]; bl = pb[s - st]; cl = pc[s - st]; dl = pd[s - st]; }
            if (s + st < ts) { au = pa[s + st]; bu = pb[s + st]; cu = pc[s + st]; du = pd[s + st]; }
            const double alpha = -a / bl;
            const double gamma = -c / bu;
            b = b + alpha * cl + gamma * au;
            d = d + alpha * dl + gamma * du;
            a = alpha * al;
            c = gamma * cu;
        }
        __syncthreads();
    }
    return d / b;
}

// Operator constants of the first row (q = 0) and of the segment a thread
// owns, loaded before the operand is known so that their memory latency
// overlaps the operand update: CSR extents and the first entries of the row's
// mB/mC couplings, P' and rfDiff coefficients, the tridiagonal JJ' entries of
// the segment, and the first kApplyChunk columns of the row of the block
// inverse.  Rows q >= 1 (tiles above kTileThreads unknowns) load in place.
constexpr int kApplyChunk = 32;
constexpr int kPreB = 8;
constexpr int kPreC = 4;

template <int D, typename TB>
struct TilePre {
    uint32_t b0, b1, c0, c1;
    uint32_t bcol[kPreB], ccol[kPreC];
    double bval[kPreB], cval[kPreC];
    double pd0, pd1, df0, df1;
    double sdf[D], sjd, sja, sjc;
    int sk, sp;
    TB blk[kApplyChunk];
};

template <int D, typename TB>
__device__ __forceinline__ void tile_prefetch(const FiberArgs& a, const TileLds<D>& L, const Tile& T,
                                              const TileRows<D>& rw, const TB* __restrict__ B, TilePre<D, TB>& pre)
{
    pre.b0 = pre.b1 = pre.c0 = pre.c1 = 0;
    pre.pd0 = pre.pd1 = pre.df0 = pre.df1 = 0.0;
#pragma unroll
    for (int v = 0; v < kPreB; ++v) { pre.bcol[v] = 0; pre.bval[v] = 0.0; }
#pragma unroll
    for (int v = 0; v < kPreC; ++v) { pre.ccol[v] = 0; pre.cval[v] = 0.0; }
#pragma unroll
    for (int v = 0; v < kApplyChunk; ++v) pre.blk[v] = TB(0);
    const int k = rw.k[0];
    if (k >= 0) {
        const int u = threadIdx.x;
        const int p = rw.p[0];
        const int fo = (int)L.fo[k];
        const int S = (int)L.fo[k + 1] - fo - 1;
        const int d = u - D * (fo + p);
        const uint32_t g = L.fseg[k] + p;
        if (B != nullptr && L.fuse[k]) {
            const int bs = D * (S + 1);
            const TB* col = B + L.fblk[k] + (u - D * fo);
#pragma unroll
            for (int v = 0; v < kApplyChunk; ++v)
                if (v < bs) pre.blk[v] = col[(size_t)bs * v];
        }
        if (a.Buse) {
            const uint32_t r = T.P0 + fo + p;
            pre.b0 = a.Brow[r];
            pre.b1 = a.Brow[r + 1];
#pragma unroll
            for (int v = 0; v < kPreB; ++v)
                if (pre.b0 + v < pre.b1) {
                    pre.bcol[v] = a.Bcol[pre.b0 + v];
                    pre.bval[v] = a.Bval[pre.b0 + v];
                }
        }
        if (a.Cuse) {
            const uint32_t r = T.U0 + u;
            pre.c0 = a.Crow[r];
            pre.c1 = a.Crow[r + 1];
#pragma unroll
            for (int v = 0; v < kPreC; ++v)
                if (pre.c0 + v < pre.c1) {
                    pre.ccol[v] = a.Ccol[pre.c0 + v];
                    pre.cval[v] = a.Cval[pre.c0 + v];
                }
        }
        if (p < S) {
            pre.pd0 = a.pdiff[g];
            pre.df0 = a.diff[D * g + d];
        }
        if (p >= 1) {
            pre.pd1 = a.pdiff[g - 1];
            pre.df1 = a.diff[D * (g - 1) + d];
        }
    }
    const int s = threadIdx.x;
#pragma unroll
    for (int d = 0; d < D; ++d) pre.sdf[d] = 0.0;
    pre.sjd = 1.0;
    pre.sja = pre.sjc = 0.0;
    pre.sk = pre.sp = 0;
    if (s < (int)T.ts) {
        const int ks = tile_fiber_of_seg(L.fo, (int)T.nft, (uint32_t)s);
        const int p = s - ((int)L.fo[ks] - ks);
        const int S = (int)L.fo[ks + 1] - (int)L.fo[ks] - 1;
        const uint32_t g = L.fseg[ks] + p;
#pragma unroll
        for (int d = 0; d < D; ++d) pre.sdf[d] = a.diff[D * g + d];
        pre.sjd = a.jd0[g];
        if (p > 0) pre.sja = a.je0[g - 1];
        if (p + 1 < S) pre.sjc = a.je0[g];
        pre.sk = ks;
        pre.sp = p;
    }
}

// sum over CSR entries [e, e1) of val * X(col * mul + add), in entry order;
// U entries' loads are issued before any is used
template <int U, typename Op>
__device__ __forceinline__ double csr_row(const uint32_t* __restrict__ col, const double* __restrict__ val,
                                          uint32_t e, uint32_t e1, const Op& X, uint32_t mul, uint32_t add,
                                          double acc)
{
    for (; e < e1; e += U) {
        uint32_t cc[U];
        double vv[U], xx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = e + u < e1;
            cc[u] = ok ? col[e + u] : 0u;
            vv[u] = ok ? val[e + u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xx[u] = (e + u < e1) ? X(cc[u] * mul + add) : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (e + u < e1) acc += vv[u] * xx[u];
    }
    return acc;
}

// the same sum with the first U entries already in registers
template <int U, typename Op>
__device__ __forceinline__ double csr_row_pre(const uint32_t* cc, const double* vv, const uint32_t* __restrict__ col,
                                              const double* __restrict__ val, uint32_t e, uint32_t e1, const Op& X,
                                              uint32_t mul, uint32_t add, double acc)
{
    double xx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xx[u] = (e + u < e1) ? X(cc[u] * mul + add) : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (e + u < e1) acc += vv[u] * xx[u];
    if (e + U < e1) acc = csr_row<U>(col, val, e + U, e1, X, mul, add, acc);
    return acc;
}

// L.y = X - dt*mob*P (R + B + C + P') X over the fibers of the tile
// (Meca::multiply, meca.cc:157-181; projection rigid_fiber_project.cc:95-149).
// L.x must hold the tile's slice of the operand and be visible to all threads;
// X(j) gives any entry of the operand (coupling gathers).  Ends synchronised.
template <int D, typename Op, typename TB>
__device__ __forceinline__ void tile_operator(const FiberArgs& a, TileLds<D>& L, const Op& X, const Tile& T,
                                              const TileRows<D>& rw, const TilePre<D, TB>& pre)
{
    constexpr int R = TileRows<D>::R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int k = rw.k[q];
        if (k < 0) continue;
        const int u = threadIdx.x + q * kTileThreads;
        const int p = rw.p[q];
        const int fo = (int)L.fo[k];
        const int np = (int)L.fo[k + 1] - fo;
        const int d = u - D * (fo + p);
        const int S = np - 1;
        double fp = 0.0;
        double pd0, pd1;
        if (q == 0) {
            if (a.Buse)
                fp = csr_row_pre<kPreB>(pre.bcol, pre.bval, a.Bcol, a.Bval, pre.b0, pre.b1, X, (uint32_t)D,
                                        (uint32_t)d, fp);
            if (a.Cuse)
                fp += csr_row_pre<kPreC>(pre.ccol, pre.cval, a.Ccol, a.Cval, pre.c0, pre.c1, X, 1u, 0u, 0.0);
            pd0 = pre.pd0;
            pd1 = pre.pd1;
        } else {
            if (a.Buse) {
                const uint32_t r = T.P0 + fo + p;
                fp = csr_row<kPreB>(a.Bcol, a.Bval, a.Brow[r], a.Brow[r + 1], X, (uint32_t)D, (uint32_t)d, fp);
            }
            if (a.Cuse) {
                const uint32_t r = T.U0 + u;
                fp += csr_row<kPreC>(a.Ccol, a.Cval, a.Crow[r], a.Crow[r + 1], X, 1u, 0u, 0.0);
            }
            const uint32_t g = L.fseg[k] + p;
            pd0 = p < S ? a.pdiff[g] : 0.0;
            pd1 = p >= 1 ? a.pdiff[g - 1] : 0.0;
        }
        const double* xf = L.x + D * fo;
        if (np > 2) {
            // bending stencil v_j = k (x_j - 2 x_{j+1} + x_{j+2}); Y_j -= v, Y_{j+1} += 2v, Y_{j+2} -= v
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
        // P': w = c_s (X_{s+1} - X_s); Y_s += w; Y_{s+1} -= w
        if (p < S && pd0 != 0.0) fp += pd0 * (xf[D * (p + 1) + d] - xf[D * p + d]);
        if (p >= 1 && pd1 != 0.0) fp -= pd1 * (xf[D * p + d] - xf[D * (p - 1) + d]);
        L.f[u] = fp;
    }
    __syncthreads();

    // multipliers: (J J') lam = J f per fiber, all fibers of the tile at once
    const int ts = (int)T.ts;
    const int s = threadIdx.x;
    double t = 0.0;
    if (s < ts) {
        const int pt = (int)L.fo[pre.sk] + pre.sp;
#pragma unroll
        for (int d = 0; d < D; ++d) t += pre.sdf[d] * (L.f[D * (pt + 1) + d] - L.f[D * pt + d]);
    }
    if (ts <= kWave) {
        if (s < kWave) {
            const double lam = pcr_solve_span(ts, L.maxS, pre.sja, pre.sjd, pre.sjc, t);
            if (s < ts) L.lam[s] = lam;
        }
    } else {
        const double lam = pcr_lds(L.pa, L.pb, L.pc, L.lam, ts, L.maxS, pre.sja, pre.sjd, pre.sjc, t);
        __syncthreads();
        if (s < ts) L.lam[s] = lam;
    }
    __syncthreads();

    // Y = X + sca * (f + J' lam)
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int k = rw.k[q];
        if (k < 0) continue;
        const int u = threadIdx.x + q * kTileThreads;
        const int p = rw.p[q];
        const int fo = (int)L.fo[k];
        const int S = (int)L.fo[k + 1] - fo - 1;
        const int sl = fo - k + p;
        const int d = u - D * (fo + p);
        double df0, df1;
        if (q == 0) {
            df0 = pre.df0;
            df1 = pre.df1;
        } else {
            const uint32_t g = L.fseg[k] + p;
            df0 = p < S ? a.diff[D * g + d] : 0.0;
            df1 = p >= 1 ? a.diff[D * (g - 1) + d] : 0.0;
        }
        double v = L.f[u];
        if (p < S) v = v + df0 * L.lam[sl];
        if (p >= 1) v = v - df1 * L.lam[sl - 1];
        const double yv = L.fsca[k] * v;
        L.y[u] = yv + L.x[u];
    }
    __syncthreads();
}

// out = Binv xin for every fiber of the tile (blas_xgemv on a column-major
// block, meca.cc:458; copy when the block is not usable, meca.cc:463).
// xin: the tile's slice in LDS; out: the tile's slice in HBM.  TB is the
// storage type of the blocks; operand and accumulation are FP64.  Row q = 0
// starts from its prefetched columns; every chunk of kApplyChunk columns is
// loaded before it is used.
template <typename TB, int D>
__device__ __forceinline__ void tile_block_apply(const TB* __restrict__ B, const TileLds<D>& L,
                                                 const TileRows<D>& rw, const TilePre<D, TB>& pre,
                                                 const double* xin, double* __restrict__ out, int stride = 1)
{
    constexpr int R = TileRows<D>::R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int k = rw.k[q];
        if (k < 0) continue;
        const int u = threadIdx.x + q * kTileThreads;
        const int fo = (int)L.fo[k];
        const int bs = D * ((int)L.fo[k + 1] - fo);
        const int i = u - D * fo;
        const double* xs = xin + D * fo;
        if (B == nullptr || !L.fuse[k]) {
            out[(size_t)u * stride] = xs[i];
            continue;
        }
        const TB* col = B + L.fblk[k] + i;
        double acc[4] = { 0, 0, 0, 0 };
        int j = 0;
        if (q == 0) {
#pragma unroll
            for (int v = 0; v < kApplyChunk; ++v)
                if (v < bs) acc[v & 3] += (double)pre.blk[v] * xs[v];
            j = kApplyChunk;
        }
        for (; j < bs; j += kApplyChunk) {
            TB c[kApplyChunk];
#pragma unroll
            for (int v = 0; v < kApplyChunk; ++v) c[v] = (j + v < bs) ? col[(size_t)bs * (j + v)] : TB(0);
#pragma unroll
            for (int v = 0; v < kApplyChunk; ++v)
                if (j + v < bs) acc[v & 3] += (double)c[v] * xs[j + v];
        }
        out[(size_t)u * stride] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
}

}  // namespace cymg
