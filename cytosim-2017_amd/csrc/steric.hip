// Steric interactions on the device: Simul::setStericInteractions
// (simul_solve.cc:139-188) with PointGrid (point_grid.cc).
//
// Items are the segments of the steric fibers ("loci") and the points of the
// steric Solids whose radius exceeds REAL_EPSILON, each filed in the grid cell
// that holds its centre (Grid::index with the bounded image, grid.h:147-150,
// 445-455), in registration order within a cell (a stable radix sort by cell).
// The reference's traversal (PointGrid::setInteractions, point_grid.cc:361-413)
// visits, cell by cell, the pairs inside the cell, then those with each cell of
// its side region (Grid::createSideRegions(1): the shifts of
// makeRectangularGrid in order, kept when inside and at an index offset >= 0).
// Here each (cell, region slot) first counts its candidate pairs; a scan gives
// every pair a rank in the reference's order, and one thread per pair runs
// checkPP / checkPL / checkLL with the reference's arithmetic.  A first pass
// counts the calls each pair makes (and their stream entries), scans place
// them, and a second pass writes them as deferred calls (interLongLink PE-PE,
// interSideSlidingLink PI-PE and PI-PI) that the deferred emitter evaluates
// from the points in that order.  Calls
// that contribute nothing in Meca (points that share a model point) are not
// made.  Where the reference orders two segments by the address of their
// records (`&bb < &aa`, point_grid.cc:205,295) their rank in the cell-major
// order decides.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "context.h"

namespace cymg {

#pragma clang fp contract(off)

namespace {

constexpr double kStEps = 128.0 * 2.220446049250313e-16;   // REAL_EPSILON

struct StArgs {
    int gdim[3];
    double inf[3], delta[3];
    uint32_t ncells;
    const double* X;
    const uint32_t* lpt;   // first point of each locus (segment)
    const uint32_t* lfib;  // its Mecable
    const uint32_t* lflag; // 1: first segment, 2: last segment
    const double* lcut;    // FiberLocus::len() = segmentation
    const double* lrad;    // FatLocus::radius
    const double* lrng;    // FatLocus::range
    const uint32_t* ppt;   // point of each steric sphere
    const uint32_t* pmec;  // its Mecable
    const double* prad;
    const uint32_t* lstart;  // loci of cell c: llist[lstart[c] .. lstart[c+1])
    const uint32_t* llist;
    const uint32_t* pstart;
    const uint32_t* plist;
    const uint32_t* lrank;   // cell-major rank of each locus
    double push, pull;
    uint32_t* cnt;           // count pass: calls, C entries, vBAS entries per pair
    uint32_t* cntC;
    uint32_t* cntS;
    const uint32_t* ocnt;    // emit pass: their exclusive scans
    const uint32_t* ocntC;
    const uint32_t* ocntS;
    DeferredCall* calls;
};

template <int D>
__device__ __forceinline__ double sdot(const double* a, const double* b)
{
    double s = a[0] * b[0];
#pragma unroll
    for (int d = 1; d < D; ++d) s = s + a[d] * b[d];
    return s;
}

// where the calls of one pair go: counted, or written from the pair's offsets
template <int D, bool EMIT>
struct Sink {
    const StArgs& a;
    uint32_t n = 0, nc = 0, ns = 0;
    uint32_t base = 0, baseC = 0, baseS = 0;
    __device__ Sink(const StArgs& args, uint32_t pair) : a(args)
    {
        if (EMIT) { base = a.ocnt[pair]; baseC = a.ocntC[pair]; baseS = a.ocntS[pair]; }
    }
    __device__ void put(uint32_t kind, uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3, double ca, double cb,
                        double len, double w)
    {
        uint32_t s[3];
        deferred_slots(kind, D, s);
        if (EMIT) {
            DeferredCall q;
            q.kind = kind;
            q.p[0] = p0; q.p[1] = p1; q.p[2] = p2; q.p[3] = p3;
            q.off[0] = 0; q.off[1] = baseC + nc; q.off[2] = baseS + ns;
            q.ca = ca; q.cb = cb; q.len = len; q.w = w;
            q.g[0] = q.g[1] = q.g[2] = 0.0;
            a.calls[base + n] = q;
        }
        ++n;
        nc += s[1];
        ns += s[2];
    }
    // interLongLink(PointExact, PointExact): nothing between neighbours (point_exact.cc:43-47)
    __device__ void long_link(uint32_t pa, uint32_t ma, uint32_t pb, uint32_t mb, double len, double w)
    {
        if (ma == mb && (pa == pb || pa == pb + 1 || pa + 1 == pb)) return;
        put(kDefLongLinkEE, pa, pb, 0, 0, 0.0, 0.0, len, w);
    }
    // interSideSlidingLink(PointInterpolated, PointExact): nothing when overlapping
    __device__ void side_sliding_ie(uint32_t a1, double ca, uint32_t b, double len, double w)
    {
        if (b == a1 || b == a1 + 1) return;
        put(kDefSideSlidingIE, a1, a1 + 1, b, 0, ca, 0.0, len, w);
    }
    __device__ void side_sliding_ii(uint32_t a1, double ca, uint32_t b1, double cb, double len, double w)
    {
        if (a1 == b1 || a1 == b1 + 1 || a1 + 1 == b1) return;
        put(kDefSideSlidingII, a1, a1 + 1, b1, b1 + 1, ca, cb, len, w);
    }
};

template <int D>
struct Walk {
    const StArgs& a;
    __device__ const double* pos(uint32_t p) const { return a.X + (size_t)D * p; }

    // FiberLocus::projectPoint0 (fiber_locus.cc:25-49)
    __device__ void project0(uint32_t l, const double* w, double& abs_, double& dis) const
    {
        const double *p1 = pos(a.lpt[l]), *p2 = pos(a.lpt[l] + 1);
        double dx[D], aw[D];
#pragma unroll
        for (int d = 0; d < D; ++d) { dx[d] = p2[d] - p1[d]; aw[d] = w[d] - p1[d]; }
        const double ls = a.lcut[l];
        abs_ = sdot<D>(aw, dx) / ls;
        if (0 <= abs_ && abs_ <= ls) dis = sdot<D>(aw, aw) - abs_ * abs_;
    }

    // checkPP (point_grid.cc:103-118)
    template <typename S>
    __device__ void pp(S& out, uint32_t pa, uint32_t ma, double ra, uint32_t pb, uint32_t mb, double rb) const
    {
        const double len = ra + rb;
        const double *wa = pos(pa), *wb = pos(pb);
        double vab[D];
#pragma unroll
        for (int d = 0; d < D; ++d) vab[d] = wb[d] - wa[d];
        if (sdot<D>(vab, vab) < len * len) out.long_link(pa, ma, pb, mb, len, a.push);
    }

    // checkPL (point_grid.cc:125-172)
    template <typename S>
    __device__ void pl(S& out, uint32_t i, uint32_t l) const
    {
        const uint32_t ap = a.ppt[i], am = a.pmec[i];
        const double arad = a.prad[i];
        const double* apos = pos(ap);
        const uint32_t lp = a.lpt[l], lm = a.lfib[l], fl = a.lflag[l];
        const double lrad = a.lrad[l], cut = a.lcut[l];
        const double len = arad + lrad;
        double ab = 0.0, d = 0.0;
        project0(l, apos, ab, d);
        if (0 <= ab) {
            if (ab <= cut) {
                if (d < len * len) out.side_sliding_ie(lp, ab / cut, ap, len, a.push);
            } else if (fl & 2) {
                pp(out, ap, am, arad, lp + 1, lm, lrad);
            }
        } else {
            if (fl & 1) {
                pp(out, ap, am, arad, lp, lm, lrad);
            } else {
                const double *q1 = pos(lp), *q0 = pos(lp - 1);
                double vab[D], dp[D];
#pragma unroll
                for (int k = 0; k < D; ++k) { vab[k] = apos[k] - q1[k]; dp[k] = q1[k] - q0[k]; }
                if (sdot<D>(vab, dp) >= 0) {
                    if (sdot<D>(vab, vab) < len * len) out.long_link(ap, am, lp, lm, len, a.push);
                }
            }
        }
    }

    // checkLL1 (point_grid.cc:178-242)
    template <typename S>
    __device__ void ll1(S& out, uint32_t la, uint32_t lb) const
    {
        const double ran = a.lrng[la] + a.lrad[lb];
        double ab = 0.0, d = INFINITY;
        const uint32_t ap = a.lpt[la], bp = a.lpt[lb];
        const double* b1 = pos(bp);
        project0(la, b1, ab, d);
        if (d < ran * ran) {
            const double len = a.lrad[la] + a.lrad[lb];
            out.side_sliding_ie(ap, ab / a.lcut[la], bp, len, d > len * len ? a.pull : a.push);
        } else if (ab < 0) {
            const double* a1 = pos(ap);
            if (a.lflag[la] & 1) {
                if (a.lrank[lb] < a.lrank[la] && (a.lflag[lb] & 1)) {
                    const double* b2 = pos(bp + 1);
                    double vab[D], db[D];
#pragma unroll
                    for (int k = 0; k < D; ++k) { vab[k] = b1[k] - a1[k]; db[k] = b2[k] - b1[k]; }
                    const double len = a.lrad[la] + a.lrad[lb];
                    if (sdot<D>(vab, vab) < len * len && sdot<D>(vab, db) >= 0)
                        out.long_link(ap, a.lfib[la], bp, a.lfib[lb], len, a.push);
                }
            } else {
                const double* a0 = pos(ap - 1);
                double vab[D], dp[D];
#pragma unroll
                for (int k = 0; k < D; ++k) { vab[k] = b1[k] - a1[k]; dp[k] = a1[k] - a0[k]; }
                if (sdot<D>(vab, dp) >= 0) {
                    const double dd = sdot<D>(vab, vab);
                    if (dd < ran * ran) {
                        const double len = a.lrad[la] + a.lrad[lb];
                        out.long_link(ap, a.lfib[la], bp, a.lfib[lb], len, dd > len * len ? a.pull : a.push);
                    }
                }
            }
        }
    }

    // checkLL2 (point_grid.cc:248-315)
    template <typename S>
    __device__ void ll2(S& out, uint32_t la, uint32_t lb) const
    {
        const double ran = a.lrng[la] + a.lrad[lb];
        double ab = 0.0, d = INFINITY;
        const uint32_t ap = a.lpt[la], bp = a.lpt[lb];
        const double* b2 = pos(bp + 1);
        project0(la, b2, ab, d);
        if (d < ran * ran) {
            const double len = a.lrad[la] + a.lrad[lb];
            out.side_sliding_ie(ap, ab / a.lcut[la], bp + 1, len, d > len * len ? a.pull : a.push);
        } else if (ab < 0) {
            const double* a1 = pos(ap);
            double vab[D];
#pragma unroll
            for (int k = 0; k < D; ++k) vab[k] = b2[k] - a1[k];
            if (a.lflag[la] & 1) {
                const double* b1 = pos(bp);
                double db[D];
#pragma unroll
                for (int k = 0; k < D; ++k) db[k] = b2[k] - b1[k];
                const double len = a.lrad[la] + a.lrad[lb];
                if (sdot<D>(vab, vab) < len * len && sdot<D>(vab, db) <= 0)
                    out.long_link(ap, a.lfib[la], bp + 1, a.lfib[lb], len, a.push);
            } else {
                const double* a0 = pos(ap - 1);
                double dp[D];
#pragma unroll
                for (int k = 0; k < D; ++k) dp[k] = a1[k] - a0[k];
                if (sdot<D>(vab, dp) >= 0) {
                    const double dd = sdot<D>(vab, vab);
                    if (dd < ran * ran) {
                        const double len = a.lrad[la] + a.lrad[lb];
                        out.long_link(ap, a.lfib[la], bp + 1, a.lfib[lb], len, dd > len * len ? a.pull : a.push);
                    }
                }
            }
        } else if (a.lrank[lb] < a.lrank[la] && (a.lflag[la] & 2) && ab > a.lcut[la]) {
            const double *a2 = pos(ap + 1), *b1 = pos(bp);
            double vab[D], db[D];
#pragma unroll
            for (int k = 0; k < D; ++k) { vab[k] = b2[k] - a2[k]; db[k] = b2[k] - b1[k]; }
            const double len = a.lrad[la] + a.lrad[lb];
            if (sdot<D>(vab, vab) < len * len && sdot<D>(vab, db) <= 0)
                out.long_link(ap + 1, a.lfib[la], bp + 1, a.lfib[lb], len, a.push);
        }
    }

    // FiberLocus::projectSegment (fiber_locus.cc:206-270)
    __device__ bool project_segment(uint32_t la, uint32_t lb, double& abs1, double& abs2, double& dist) const
    {
        const double *a1 = pos(a.lpt[la]), *a2 = pos(a.lpt[la] + 1);
        const double *b1 = pos(a.lpt[lb]), *b2 = pos(a.lpt[lb] + 1);
        double d1[D], d2[D], d12[D];
#pragma unroll
        for (int k = 0; k < D; ++k) { d1[k] = a2[k] - a1[k]; d2[k] = b2[k] - b1[k]; d12[k] = b1[k] - a1[k]; }
        const double len1 = a.lcut[la], len2 = a.lcut[lb];
        const double beta = sdot<D>(d1, d2) / (len1 * len2);
        const double scal = 1.0 - beta * beta;
        if (scal > kStEps) {
            const double d1d12 = sdot<D>(d1, d12) / (scal * len1);
            const double d2d12 = sdot<D>(d2, d12) / (scal * len2);
            abs1 = d1d12 - beta * d2d12;
            if (abs1 < 0 || len1 < abs1) return false;
            abs2 = beta * d1d12 - d2d12;
            if (abs2 < 0 || len2 < abs2) return false;
            const double s = abs2 / len2;
            double v[D];
#pragma unroll
            for (int k = 0; k < D; ++k) v[k] = d12[k] + d2[k] * s;
            dist = sdot<D>(v, v) - abs1 * abs1;
            return true;
        }
        double p1 = sdot<D>(d12, d1) / len1;
        double p2 = p1 + beta * len2;
        if (p1 < 0 && p2 < 0) return false;
        if (p1 > len1 && p2 > len1) return false;
        const double p = p1;
        dist = sdot<D>(d12, d12) - p1 * p1;
        if (p1 < 0) p1 = 0; else if (p1 > len1) p1 = len1;
        if (p2 < 0) p2 = 0; else if (p2 > len1) p2 = len1;
        abs1 = 0.5 * (p1 + p2);
        abs2 = (abs1 - p) * beta;
        return true;
    }

    // checkLL (point_grid.cc:321-354)
    template <typename S>
    __device__ void ll(S& out, uint32_t la, uint32_t lb) const
    {
        ll1(out, la, lb);
        if (a.lflag[la] & 2) ll2(out, lb, la);
        ll1(out, lb, la);
        if (a.lflag[lb] & 2) ll2(out, la, lb);
        if constexpr (D == 3) {
            const double ran = fmax(a.lrng[la] + a.lrad[lb], a.lrad[la] + a.lrng[lb]);
            double s1, s2, d;
            if (project_segment(la, lb, s1, s2, d) && d < ran * ran) {
                const double len = a.lrad[la] + a.lrad[lb];
                out.side_sliding_ii(a.lpt[la], s1 / a.lcut[la], a.lpt[lb], s2 / a.lcut[lb], len,
                                    d > len * len ? a.pull : a.push);
            }
        }
    }
};

// the side region of a cell (Grid::createSideRegions(1), grid.h:650-861):
// offset of its r-th cell (r = 0: the cell itself), or -1 past its last;
// r < 0 counts its cells
template <int D>
__device__ int side_region(const StArgs& a, uint32_t indx, int r)
{
    int ori[3] = { 0, 0, 0 };
    uint32_t rem = indx;
    for (int d = 0; d < D; ++d) { ori[d] = (int)(rem % (uint32_t)a.gdim[d]); rem /= (uint32_t)a.gdim[d]; }
    int cnt = 0;
    constexpr int NS = D == 3 ? 27 : 9;
    for (int q = 0; q < NS; ++q) {
        // shift q of makeRectangularGrid: per dimension d the digit (q / 3^d) % 3
        // in the order 0, -1, +1
        int sh[3] = { 0, 0, 0 }, t = q;
        for (int d = 0; d < D; ++d) {
            const int dg = t % 3;
            t /= 3;
            sh[d] = dg == 0 ? 0 : (dg == 1 ? -1 : 1);
        }
        int cc[3], idx = 0;
        bool in = true;
        for (int d = 0; d < D; ++d) {
            cc[d] = ori[d] + sh[d];
            if (cc[d] < 0 || cc[d] >= a.gdim[d]) in = false;
        }
        if (!in) continue;
        for (int d = D - 1; d >= 0; --d) idx = a.gdim[d] * idx + cc[d];
        const int o = idx - (int)indx;
        if (o < 0) continue;
        if (cnt == r) return o;
        ++cnt;
    }
    return r < 0 ? cnt : -1;
}

// (cell, region slot) pairs: slot 0 the pairs inside the cell, slot r >= 1
// the pairs with the r-th cell of its side region; the slots of a cell follow
// each other, so ranks in cell-major order keep the reference's order
template <int D>
constexpr int kSlots = D == 3 ? 14 : 5;   // (3^D + 1) / 2 cells in a side region

// candidate pairs of each slot: inside a cell the n(n-1)/2 pairs of its items
// (points first, then loci), with a neighbour cell its items times theirs
template <int D>
__global__ __launch_bounds__(256) void k_st_tasks(StArgs a, uint64_t* __restrict__ ntask)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= a.ncells * (uint32_t)kSlots<D>) return;
    const uint32_t indx = u / kSlots<D>;
    const int r = (int)(u % kSlots<D>);
    const uint64_t n = a.pstart[indx + 1] - a.pstart[indx] + a.lstart[indx + 1] - a.lstart[indx];
    uint64_t m = 0;
    if (n) {
        if (r == 0) {
            m = n * (n - 1) / 2;
        } else {
            const int o = side_region<D>(a, indx, r);
            if (o > 0) {
                const uint32_t side = indx + (uint32_t)o;
                m = n * (a.pstart[side + 1] - a.pstart[side] + a.lstart[side + 1] - a.lstart[side]);
            }
        }
    }
    ntask[u] = m;
}

// the slot of every pair
__global__ __launch_bounds__(256) void k_st_fill(uint32_t nw, const uint64_t* __restrict__ toff,
                                                 uint32_t* __restrict__ slot)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nw) return;
    const uint64_t e = toff[u + 1];
    for (uint64_t t = toff[u]; t < e; ++t) slot[t] = u;
}

// one thread per candidate pair, in the order of the reference's loops: inside
// a cell, for each item (points, then loci) the later items; with a neighbour
// cell, for each item of the cell the neighbour's points, then its loci
template <int D, bool EMIT>
__global__ __launch_bounds__(256) void k_st_pairs(StArgs a, uint32_t npair, const uint64_t* __restrict__ toff,
                                                  const uint32_t* __restrict__ slot)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npair) return;
    const uint32_t u = slot[t];
    uint32_t k = (uint32_t)(t - toff[u]);
    const uint32_t indx = u / kSlots<D>;
    const int r = (int)(u % kSlots<D>);
    Sink<D, EMIT> out(a, t);
    const Walk<D> W{ a };
    const uint32_t p0 = a.pstart[indx], P = a.pstart[indx + 1] - p0;
    const uint32_t l0 = a.lstart[indx];
    if (r == 0) {
        const uint32_t n = P + a.lstart[indx + 1] - l0;
        uint32_t ii = 0;
        while (k >= n - 1 - ii) { k -= n - 1 - ii; ++ii; }
        const uint32_t jj = ii + 1 + k;
        if (jj < P) {
            const uint32_t i = a.plist[p0 + ii], j = a.plist[p0 + jj];
            W.pp(out, a.ppt[i], a.pmec[i], a.prad[i], a.ppt[j], a.pmec[j], a.prad[j]);
        } else if (ii < P) {
            W.pl(out, a.plist[p0 + ii], a.llist[l0 + jj - P]);
        } else {
            W.ll(out, a.llist[l0 + ii - P], a.llist[l0 + jj - P]);
        }
    } else {
        const uint32_t side = indx + (uint32_t)side_region<D>(a, indx, r);
        const uint32_t q0 = a.pstart[side], Q = a.pstart[side + 1] - q0;
        const uint32_t s0 = a.lstart[side], n2 = Q + a.lstart[side + 1] - s0;
        const uint32_t row = k / n2, col = k - row * n2;
        if (row < P) {
            const uint32_t i = a.plist[p0 + row];
            if (col < Q) {
                const uint32_t j = a.plist[q0 + col];
                W.pp(out, a.ppt[i], a.pmec[i], a.prad[i], a.ppt[j], a.pmec[j], a.prad[j]);
            } else {
                W.pl(out, i, a.llist[s0 + col - Q]);
            }
        } else {
            const uint32_t li = a.llist[l0 + row - P];
            if (col < Q) W.pl(out, a.plist[q0 + col], li);
            else W.ll(out, li, a.llist[s0 + col - Q]);
        }
    }
    if (!EMIT) {
        a.cnt[t] = out.n;
        a.cntC[t] = out.nc;
        a.cntS[t] = out.ns;
    }
}

// cell of each item (Grid::index, bounded image)
template <int D>
__device__ uint32_t cell_of(const StArgs& a, const double* w)
{
    uint32_t inx = 0;
    for (int d = D - 1; d >= 0; --d) {
        const double c = 0.0 + (w[d] - a.inf[d]) * a.delta[d];
        const int s = a.gdim[d];
        const int i = c <= 0 ? 0 : (c >= s ? s - 1 : (int)c);
        inx = (uint32_t)a.gdim[d] * inx + (uint32_t)i;
    }
    return inx;
}

template <int D>
__global__ void k_st_cells(StArgs a, uint32_t nloc, uint32_t npnt, uint32_t* lkey, uint32_t* lval, uint32_t* pkey,
                           uint32_t* pval)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nloc) {
        const double *p1 = a.X + (size_t)D * a.lpt[t], *p2 = p1 + D;
        double c[D];
#pragma unroll
        for (int d = 0; d < D; ++d) c[d] = 0.5 * (p1[d] + p2[d]);
        lkey[t] = cell_of<D>(a, c);
        lval[t] = t;
    } else if (t < nloc + npnt) {
        const uint32_t i = t - nloc;
        pkey[i] = cell_of<D>(a, a.X + (size_t)D * a.ppt[i]);
        pval[i] = i;
    }
}

// start of each cell's run in a sorted key array (ncells + 1 entries)
__global__ void k_st_starts(uint32_t ncells, uint32_t n, const uint32_t* __restrict__ key, uint32_t* __restrict__ start)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key[mid] < c) lo = mid + 1; else hi = mid;
    }
    start[c] = lo;
}

__global__ void k_st_ranks(uint32_t n, const uint32_t* __restrict__ sorted_item, uint32_t* __restrict__ rank)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) rank[sorted_item[k]] = k;
}

static inline dim3 blocks_for(size_t n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

}  // namespace

// Items of the steric Mecables of the current layout and the grid geometry
// (PointGrid::setGrid, point_grid.cc:25-64; PointGrid::add range checks,
// point_grid.cc:71-96); made again when the layout changes.
static void steric_items(Context& c)
{
    if (c.st_gen == c.layout_gen) return;
    const int D = c.dim;
    const uint32_t nmec = c.nmec();
    if (c.h_st_flag.size() != nmec)
        throw Error(CYMG_ERR_INVALID, "cymg_set_steric must describe every Mecable registered");
    std::vector<uint32_t> lpt, lfib, lflag, ppt, pmec;
    std::vector<double> lcut, lrad, lrng, prad;
    // Simul::estimateStericRange (simul_solve.cc:12-82), with the fibers' segmentation
    double step = c.st_prm.max_range;
    if (step <= 0) {
        double ran = 0, len = 0;
        for (uint32_t f = 0; f < c.nfib; ++f) {
            if (!c.h_st_flag[f]) continue;
            len = std::max(len, 4.1 / 3 * c.h_cut[f]);
            ran = std::max(ran, c.h_st_rad[f] + c.h_st_rng[f]);
        }
        ran = len + 2 * ran;
        for (uint32_t k = 0; k < c.nsol; ++k) {
            if (!c.h_st_flag[c.nfib + k]) continue;
            const uint32_t b = c.h_sol_idx[k] - c.fib_pts;
            for (uint32_t p = 0; p < c.h_sol_np[k]; ++p) ran = std::max(ran, 2 * c.h_sol_rad[b + p]);
        }
        step = ran;
    }
    if (!(step > kStEps)) throw Error(CYMG_ERR_INVALID, "simul:steric is enabled, but simul:steric_max_range was not set");
    double width[3] = { 0, 0, 0 };
    uint64_t ncells = 1;
    for (int d = 0; d < D; ++d) {
        double ext = c.st_prm.extension[d];
        if (ext < 0) throw Error(CYMG_ERR_INVALID, "space:dimension should be >= 0");
        const double n = 2 * ext / step;
        c.st_gdim[d] = (int)std::ceil(n) + 2;
        ext = c.st_gdim[d] * 0.5 * step;
        c.st_inf[d] = -ext;
        width[d] = (ext - c.st_inf[d]) / (double)c.st_gdim[d];
        c.st_delta[d] = (double)c.st_gdim[d] / (ext - c.st_inf[d]);
        ncells *= (uint64_t)c.st_gdim[d];
    }
    if (ncells > (1ull << 23)) throw Error(CYMG_ERR_UNSUPPORTED, "steric grid of more than 2^23 cells");
    c.st_ncells = (uint32_t)ncells;
    double maxd = 0;   // Grid::minimumWidth(1), grid.h:354-364
    for (int d = 0; d < D; ++d) if (c.st_gdim[d] > 1) maxd = width[d];
    for (int d = 0; d < D; ++d) if (c.st_gdim[d] > 1 && width[d] < maxd) maxd = width[d];
    for (uint32_t f = 0; f < c.nfib; ++f) {
        if (!c.h_st_flag[f]) continue;
        const double cut = c.h_cut[f], rad = c.h_st_rad[f], rng = rad + c.h_st_rng[f];
        if (maxd < std::sqrt(cut * cut + rng * rng))
            throw Error(CYMG_ERR_INVALID, "simul:steric_max_range is too short: steric_max_range should be greater "
                                          "than sqrt( sqr(segment_length) + 4*sqr(range) )");
        for (uint32_t r = 0; r + 1 < c.h_np[f]; ++r) {
            lpt.push_back(c.h_idx[f] + r);
            lfib.push_back(f);
            lflag.push_back((r == 0 ? 1u : 0u) | (r + 2 == c.h_np[f] ? 2u : 0u));
            lcut.push_back(cut);
            lrad.push_back(rad);
            lrng.push_back(rng);
        }
    }
    for (uint32_t k = 0; k < c.nsol; ++k) {
        if (!c.h_st_flag[c.nfib + k]) continue;
        const uint32_t b = c.h_sol_idx[k] - c.fib_pts;
        for (uint32_t p = 0; p < c.h_sol_np[k]; ++p) {
            const double r = c.h_sol_rad[b + p];
            if (!(r > kStEps)) continue;
            if (maxd < 2 * r)
                throw Error(CYMG_ERR_INVALID, "simul:steric_max_range is too short: steric_max_range should be "
                                              "greater than 2 * particle-radius");
            ppt.push_back(c.h_sol_idx[k] + p);
            pmec.push_back(c.nfib + k);
            prad.push_back(r);
        }
    }
    c.st_nloc = (uint32_t)lpt.size();
    c.st_npnt = (uint32_t)ppt.size();
    auto up = [&](auto& dst, const auto& src) {
        dst.resize(src.size() ? src.size() : 1);
        if (!src.empty())
            CYMG_HIP(hipMemcpyAsync(dst.get(), src.data(), src.size() * sizeof(src[0]), hipMemcpyHostToDevice,
                                    c.stream));
    };
    up(c.st_lpt, lpt); up(c.st_lfib, lfib); up(c.st_lflag, lflag);
    up(c.st_lcut, lcut); up(c.st_lrad, lrad); up(c.st_lrng, lrng);
    up(c.st_ppt, ppt); up(c.st_pmec, pmec); up(c.st_prad, prad);
    CYMG_HIP(hipStreamSynchronize(c.stream));   // pageable sources
    c.st_gen = c.layout_gen;
}

// Simul::setStericInteractions: the calls of the grid in its order, as one
// device emitter mark; returns after one device-to-host copy of their number
void steric_interactions(Context& c)
{
    steric_items(c);
    const int D = c.dim;
    const uint32_t nl = c.st_nloc, np = c.st_npnt, nc = c.st_ncells;
    StArgs a;
    for (int d = 0; d < 3; ++d) { a.gdim[d] = c.st_gdim[d]; a.inf[d] = c.st_inf[d]; a.delta[d] = c.st_delta[d]; }
    a.ncells = nc;
    a.X = c.d_pos.get();
    a.lpt = c.st_lpt.get(); a.lfib = c.st_lfib.get(); a.lflag = c.st_lflag.get();
    a.lcut = c.st_lcut.get(); a.lrad = c.st_lrad.get(); a.lrng = c.st_lrng.get();
    a.ppt = c.st_ppt.get(); a.pmec = c.st_pmec.get(); a.prad = c.st_prad.get();
    a.push = c.st_prm.stiff_push;
    a.pull = c.st_prm.stiff_pull;
    const size_t ni = (size_t)nl + np;
    c.st_key.resize(ni ? ni : 1); c.st_val.resize(ni ? ni : 1);
    c.st_key2.resize(ni ? ni : 1); c.st_val2.resize(ni ? ni : 1);
    uint32_t *lkey = c.st_key.get(), *lval = c.st_val.get(), *pkey = lkey + nl, *pval = lval + nl;
    uint32_t *lkey2 = c.st_key2.get(), *lval2 = c.st_val2.get(), *pkey2 = lkey2 + nl, *pval2 = lval2 + nl;
    hipStream_t st = c.stream;
    if (ni) {
        if (D == 3) hipLaunchKernelGGL(k_st_cells<3>, blocks_for(ni, 256), dim3(256), 0, st, a, nl, np, lkey, lval, pkey, pval);
        else hipLaunchKernelGGL(k_st_cells<2>, blocks_for(ni, 256), dim3(256), 0, st, a, nl, np, lkey, lval, pkey, pval);
    }
    // stable sort by cell: registration order kept inside a cell
    int end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) <= nc) ++end_bit;
    size_t tb = 0, tb2 = 0, tb3 = 0;
    if (nl) CYMG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, lkey, lkey2, lval, lval2, (int)nl, 0, end_bit, st));
    if (np) CYMG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, pkey, pkey2, pval, pval2, (int)np, 0, end_bit, st));
    const uint32_t nw = nc * (uint32_t)(D == 3 ? kSlots<3> : kSlots<2>);   // (cell, region slot)
    c.st_ntask.resize((size_t)nw + 1);
    c.st_toff.resize((size_t)nw + 1);
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, c.st_ntask.get(), c.st_toff.get(), (int)(nw + 1), st));
    c.st_temp.resize(std::max({ tb, tb2, tb3, (size_t)1 }));
    if (nl) CYMG_HIP(hipcub::DeviceRadixSort::SortPairs(c.st_temp.get(), tb, lkey, lkey2, lval, lval2, (int)nl, 0, end_bit, st));
    if (np) CYMG_HIP(hipcub::DeviceRadixSort::SortPairs(c.st_temp.get(), tb2, pkey, pkey2, pval, pval2, (int)np, 0, end_bit, st));
    c.st_lstart.resize((size_t)nc + 1); c.st_pstart.resize((size_t)nc + 1);
    c.st_lrank.resize(nl ? nl : 1);
    hipLaunchKernelGGL(k_st_starts, blocks_for((size_t)nc + 1, 256), dim3(256), 0, st, nc, nl, lkey2, c.st_lstart.get());
    hipLaunchKernelGGL(k_st_starts, blocks_for((size_t)nc + 1, 256), dim3(256), 0, st, nc, np, pkey2, c.st_pstart.get());
    if (nl) hipLaunchKernelGGL(k_st_ranks, blocks_for(nl, 256), dim3(256), 0, st, nl, lval2, c.st_lrank.get());
    a.lstart = c.st_lstart.get(); a.llist = lval2;
    a.pstart = c.st_pstart.get(); a.plist = pval2;
    a.lrank = c.st_lrank.get();
    // candidate pairs per slot, their ranks, and their number on the host
    CYMG_HIP(hipMemsetAsync(c.st_ntask.get() + nw, 0, 8, st));
    if (D == 3) hipLaunchKernelGGL(k_st_tasks<3>, blocks_for(nw, 256), dim3(256), 0, st, a, c.st_ntask.get());
    else hipLaunchKernelGGL(k_st_tasks<2>, blocks_for(nw, 256), dim3(256), 0, st, a, c.st_ntask.get());
    CYMG_HIP(hipGetLastError());
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(c.st_temp.get(), tb3, c.st_ntask.get(), c.st_toff.get(), (int)(nw + 1), st));
    if (!c.h_st_tot) CYMG_HIP(hipHostMalloc((void**)&c.h_st_tot, 4 * sizeof(uint32_t), hipHostMallocDefault));
    uint64_t* h_npair = reinterpret_cast<uint64_t*>(c.h_st_tot + 2);
    CYMG_HIP(hipMemcpyAsync(h_npair, c.st_toff.get() + nw, 8, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipStreamSynchronize(st));
    const uint64_t npair64 = *h_npair;
    if (npair64 >= (1ull << 30)) throw Error(CYMG_ERR_UNSUPPORTED, "steric grid with more than 2^30 candidate pairs");
    const uint32_t npair = (uint32_t)npair64;
    c.st_slot.resize(npair ? npair : 1);
    c.st_cnt.resize((size_t)npair + 1); c.st_cntC.resize((size_t)npair + 1); c.st_cntS.resize((size_t)npair + 1);
    c.st_ocnt.resize((size_t)npair + 1); c.st_ocntC.resize((size_t)npair + 1); c.st_ocntS.resize((size_t)npair + 1);
    size_t tb4 = 0;
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb4, c.st_cnt.get(), c.st_ocnt.get(), (int)(npair + 1), st));
    c.st_temp.resize(std::max(c.st_temp.n, tb4));
    a.cnt = c.st_cnt.get(); a.cntC = c.st_cntC.get(); a.cntS = c.st_cntS.get();
    // the last entry stays 0, so the scans end with the totals
    CYMG_HIP(hipMemsetAsync(a.cnt + npair, 0, 4, st));
    CYMG_HIP(hipMemsetAsync(a.cntC + npair, 0, 4, st));
    CYMG_HIP(hipMemsetAsync(a.cntS + npair, 0, 4, st));
    if (npair) {
        hipLaunchKernelGGL(k_st_fill, blocks_for(nw, 256), dim3(256), 0, st, nw, c.st_toff.get(), c.st_slot.get());
        if (D == 3) hipLaunchKernelGGL((k_st_pairs<3, false>), blocks_for(npair, 256), dim3(256), 0, st, a, npair,
                                       c.st_toff.get(), c.st_slot.get());
        else hipLaunchKernelGGL((k_st_pairs<2, false>), blocks_for(npair, 256), dim3(256), 0, st, a, npair,
                                c.st_toff.get(), c.st_slot.get());
        CYMG_HIP(hipGetLastError());
    }
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(c.st_temp.get(), tb4, a.cnt, c.st_ocnt.get(), (int)(npair + 1), st));
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(c.st_temp.get(), tb4, a.cntC, c.st_ocntC.get(), (int)(npair + 1), st));
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(c.st_temp.get(), tb4, a.cntS, c.st_ocntS.get(), (int)(npair + 1), st));
    CYMG_HIP(hipMemcpyAsync(c.h_st_tot, c.st_ocnt.get() + npair, 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(c.h_st_tot + 1, c.st_ocntC.get() + npair, 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(c.h_st_tot + 2, c.st_ocntS.get() + npair, 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipStreamSynchronize(st));
    const uint32_t ncalls = c.h_st_tot[0];
    // calls of earlier steric marks of this assembly stay in front
    uint32_t first = 0;
    for (const EmitterMark& mk : c.marks)
        if (mk.kind == 5) first = std::max(first, mk.first + mk.count);
    if (ncalls) {
        DevBuf<DeferredCall> keep;
        if (first) {
            keep.resize(first);
            CYMG_HIP(hipMemcpyAsync(keep.get(), c.d_st_calls.get(), (size_t)first * sizeof(DeferredCall),
                                    hipMemcpyDeviceToDevice, st));
        }
        c.d_st_calls.resize((size_t)first + ncalls);
        if (first)
            CYMG_HIP(hipMemcpyAsync(c.d_st_calls.get(), keep.get(), (size_t)first * sizeof(DeferredCall),
                                    hipMemcpyDeviceToDevice, st));
        a.ocnt = c.st_ocnt.get(); a.ocntC = c.st_ocntC.get(); a.ocntS = c.st_ocntS.get();
        a.calls = c.d_st_calls.get() + first;
        if (D == 3) hipLaunchKernelGGL((k_st_pairs<3, true>), blocks_for(npair, 256), dim3(256), 0, st, a, npair,
                                       c.st_toff.get(), c.st_slot.get());
        else hipLaunchKernelGGL((k_st_pairs<2, true>), blocks_for(npair, 256), dim3(256), 0, st, a, npair,
                                c.st_toff.get(), c.st_slot.get());
        CYMG_HIP(hipGetLastError());
        if (first) CYMG_HIP(hipStreamSynchronize(st));   // `keep` is freed on return
    }
    EmitterMark mk{ 5, 0.0, 0.0, c.hB.key.size(), c.hC.key.size(), c.hBas.key.size() };
    mk.first = first;
    mk.count = ncalls;
    mk.slots[0] = 0;
    mk.slots[1] = c.h_st_tot[1];
    mk.slots[2] = c.h_st_tot[2];
    c.marks.push_back(mk);
}

}  // namespace cymg
