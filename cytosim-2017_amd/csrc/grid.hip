// Segments near many places at once: the query of FiberGrid (fiber_grid.cc:
// 248-337) for a batch of places on the device.
//
// The reference paints every segment, inflated by the binding range plus half
// a cell diagonal, into the cells of a fixed grid over the Space
// (fiber_grid.cc:199-241), then tests the segments listed in the place's cell
// with FiberLocus::projectPoint.  Painting is a scatter of variable length per
// segment; here each segment goes into exactly one cell, the cell of its
// midpoint, on a hashed grid whose cell side h = range + the longest
// half-segment: a segment within `range` of a place then has its midpoint in
// the place's cell or one of its neighbours (27 in 3D, 9 in 2D).  Segments are
// sorted by cell hash (one radix sort), each bucket is a run of the sorted
// list, and a place visits its neighbour cells, keeping a segment only when
// its own cell is the visited one (which removes hash collisions and
// duplicates) and projectPoint accepts it; a wavefront serves one place, a
// lane per neighbour cell.  The kept set is the reference's
// (the grid only prunes); the output is sorted by segment, which is the
// reference's order (segments are painted fiber by fiber, point by point).
#include "common.h"
#include "context.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#pragma clang fp contract(off)

namespace cymg {

constexpr int kGridHashBits = 20;
constexpr uint32_t kGridBuckets = 1u << kGridHashBits;
constexpr int kGridThreads = 256;
constexpr int kGridMaxKeep = 64;     // kept segments per place (more: error)

__device__ __forceinline__ uint32_t cell_hash(long long x, long long y, long long z)
{
    const uint64_t h = (uint64_t)x * 73856093ull ^ (uint64_t)y * 19349663ull ^ (uint64_t)z * 83492791ull;
    return (uint32_t)(h ^ (h >> 29)) & (kGridBuckets - 1);
}

template <int D>
__device__ __forceinline__ void cell_of(const double* m, double inv_h, long long c[3])
{
    c[0] = c[1] = c[2] = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) c[d] = (long long)floor(m[d] * inv_h);
}

// segment starting at fiber point P (P < fib_pts and not the fiber's last
// point): midpoint, and the largest half-length over all segments (the bits of
// a non-negative double order as the double)
template <int D>
__global__ __launch_bounds__(kGridThreads) void k_seg_half(uint32_t fib_pts, const uint32_t* __restrict__ pfib,
                                                           const uint32_t* __restrict__ idx,
                                                           const uint32_t* __restrict__ np,
                                                           const double* __restrict__ pos,
                                                           unsigned long long* __restrict__ half_bits)
{
    const uint32_t P = blockIdx.x * kGridThreads + threadIdx.x;
    double half = 0.0;
    if (P < fib_pts) {
        const uint32_t f = pfib[P];
        if (P + 1 < idx[f] + np[f]) {
            double s = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double e = pos[D * (P + 1) + d] - pos[D * P + d];
                s += e * e;
            }
            half = 0.5 * sqrt(s);
        }
    }
    for (int o = 32; o > 0; o >>= 1) half = fmax(half, __shfl_xor(half, o, 64));
    if ((threadIdx.x & 63) == 0 && half > 0.0)
        atomicMax(half_bits, (unsigned long long)__double_as_longlong(half));
}

// key (cell hash of the midpoint) and value (first point P) of every segment;
// segment s of fiber f starts at point P = s + f
template <int D>
__global__ __launch_bounds__(kGridThreads) void k_seg_keys(uint32_t fib_pts, const uint32_t* __restrict__ pfib,
                                                           const uint32_t* __restrict__ idx,
                                                           const uint32_t* __restrict__ np,
                                                           const double* __restrict__ pos, double inv_h,
                                                           uint32_t* __restrict__ key, uint32_t* __restrict__ seg)
{
    const uint32_t P = blockIdx.x * kGridThreads + threadIdx.x;
    if (P >= fib_pts) return;
    const uint32_t f = pfib[P];
    if (P + 1 >= idx[f] + np[f]) return;
    const uint32_t s = P - f;
    double m[3] = { 0, 0, 0 };
#pragma unroll
    for (int d = 0; d < D; ++d) m[d] = 0.5 * (pos[D * P + d] + pos[D * (P + 1) + d]);
    long long c[3];
    cell_of<D>(m, inv_h, c);
    key[s] = cell_hash(c[0], c[1], c[2]);
    seg[s] = P;
}

// bucket [start, end) of every hash present in the sorted keys
__global__ __launch_bounds__(kGridThreads) void k_buckets(uint32_t nseg, const uint32_t* __restrict__ key,
                                                          uint32_t* __restrict__ start, uint32_t* __restrict__ end)
{
    const uint32_t i = blockIdx.x * kGridThreads + threadIdx.x;
    if (i >= nseg) return;
    const uint32_t k = key[i];
    if (i == 0 || key[i - 1] != k) start[k] = i;
    if (i + 1 == nseg || key[i + 1] != k) end[k] = i + 1;
}

struct GridQuery {
    const double* places;      // [nq][D]
    const uint32_t* exclude;   // fiber to skip per place, or null
    double range2;
    int inclusive;
    double inv_h;
    const uint32_t* sorted;    // first points of the segments, by bucket
    const uint32_t* start;
    const uint32_t* end;
    const uint32_t* pfib;      // fiber of each point
    const uint32_t* idx;
    const uint32_t* np;
    const double* cut;
    const double* pos;
    const uint32_t* offset;    // fill pass: first output of each place
    uint32_t* count;           // count pass: kept segments of each place
    uint32_t* out_fib;
    uint32_t* out_pt;
    double* out_abs;
    double* out_dis;
    int* overflow;
};

// FiberLocus::projectPoint (fiber_locus.cc:62-97) and the range test; returns
// whether the segment is kept and its abscissa from the minus end
template <int D>
__device__ __forceinline__ bool project_keep(const GridQuery& q, const double* w, uint32_t f, uint32_t p,
                                             double& abscissa, double& dis)
{
    const double ls = q.cut[f];
    const double* P1 = q.pos + (size_t)D * (q.idx[f] + p);
    const double* P2 = P1 + D;
    double dx[3] = { 0, 0, 0 }, aw[3] = { 0, 0, 0 };
#pragma unroll
    for (int d = 0; d < D; ++d) {
        dx[d] = P2[d] - P1[d];
        aw[d] = w[d] - P1[d];
    }
    double dot = aw[0] * dx[0];
#pragma unroll
    for (int d = 1; d < D; ++d) dot = dot + aw[d] * dx[d];
    const double ab = dot / ls;
    double ds = __builtin_inf();
    if (ab < 0) {
        if (p == 0) {
            ds = (P1[0] - w[0]) * (P1[0] - w[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) ds = ds + (P1[d] - w[d]) * (P1[d] - w[d]);
        }
    } else if (ab > ls) {
        if (p + 2 == q.np[f]) {
            ds = (P2[0] - w[0]) * (P2[0] - w[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) ds = ds + (P2[d] - w[d]) * (P2[d] - w[d]);
        }
    } else {
        double nn = aw[0] * aw[0];
#pragma unroll
        for (int d = 1; d < D; ++d) nn = nn + aw[d] * aw[d];
        ds = nn - ab * ab;
    }
    abscissa = ls * (double)p + ab;
    dis = ds;
    return q.inclusive ? !(ds > q.range2) : (ds < q.range2);
}

// A wavefront per place: lane t < 27 (9 in 2D) walks neighbour cell t, so the
// cells are searched side by side; kept segments are appended to a per-wave
// LDS list, and each of them then finds its rank (the number of kept segments
// with a smaller first point) to write the output in segment order.
constexpr int kGridWaves = kGridThreads / 64;

template <int D, bool FILL>
__global__ __launch_bounds__(kGridThreads) void k_query_wave(uint32_t nq, GridQuery q)
{
    constexpr int NC = D == 3 ? 27 : 9;
    __shared__ uint32_t list[kGridWaves][kGridMaxKeep];
    __shared__ uint32_t fill[kGridWaves];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * kGridWaves + wv;
    if (lane == 0) fill[wv] = 0;
    __syncthreads();
    if (i >= nq) return;    // the whole wavefront leaves together
    double w[3] = { 0, 0, 0 };
#pragma unroll
    for (int d = 0; d < D; ++d) w[d] = q.places[(size_t)D * i + d];
    const uint32_t skip = q.exclude ? q.exclude[i] : 0xffffffffu;
    long long c[3];
    cell_of<D>(w, q.inv_h, c);
    uint32_t n = 0;
    if (lane < NC) {
        const long long cx = c[0] + (lane % 3 - 1), cy = c[1] + ((lane / 3) % 3 - 1),
                        cz = c[2] + (D == 3 ? lane / 9 - 1 : 0);
        const uint32_t h = cell_hash(cx, cy, cz);
        const uint32_t e = q.end[h];
        for (uint32_t k = e > 0 ? q.start[h] : 0; k < e; ++k) {
            const uint32_t P = q.sorted[k];
            const uint32_t f = q.pfib[P];
            if (f == skip) continue;
            double m[3] = { 0, 0, 0 };
#pragma unroll
            for (int d = 0; d < D; ++d) m[d] = 0.5 * (q.pos[D * P + d] + q.pos[D * (P + 1) + d]);
            long long sc[3];
            cell_of<D>(m, q.inv_h, sc);
            if (sc[0] != cx || sc[1] != cy || (D == 3 && sc[2] != cz)) continue;
            double ab, ds;
            if (!project_keep<D>(q, w, f, P - q.idx[f], ab, ds)) continue;
            ++n;
            if (FILL) {
                const uint32_t slot = atomicAdd(&fill[wv], 1u);
                if (slot < (uint32_t)kGridMaxKeep) list[wv][slot] = P;
                else *q.overflow = 1;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    if (!FILL) {
        if (lane == 0) q.count[i] = n;
        return;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t m = n < (uint32_t)kGridMaxKeep ? n : (uint32_t)kGridMaxKeep;
    if ((uint32_t)lane < m) {
        const uint32_t P = list[wv][lane];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < m; ++j) rank += list[wv][j] < P ? 1u : 0u;
        const uint32_t f = q.pfib[P], p = P - q.idx[f];
        double ab, ds;
        project_keep<D>(q, w, f, p, ab, ds);
        const uint32_t o = q.offset[i] + rank;
        q.out_fib[o] = f;
        q.out_pt[o] = p;
        q.out_abs[o] = ab;
        q.out_dis[o] = ds;
    }
}

// Host side of cymg_nearby_segments.  Cell side h: a kept segment has
// |w - P1|^2 - abs^2 <= range^2 with 0 <= abs <= segmentation (or is within
// range of a fiber end), so its midpoint lies within sqrt(range^2 +
// segmentation^2) + half its length of the place; h is that bound for the
// longest segmentation and segment, enlarged by 1e-6 against rounding.
void nearby_segments(Context& c, uint32_t nq, const double* places, const uint32_t* exclude, double range,
                     int inclusive, uint32_t* counts, uint32_t capacity, uint32_t* fiber, uint32_t* point,
                     double* abscissa, double* dist_sqr, uint32_t* total)
{
    hipStream_t st = c.stream;
    const int D = c.dim;
    const uint32_t fib_pts = c.fib_pts, nseg = c.nseg;
    if (!nq) {
        *total = 0;
        return;
    }
    GridBufs& G = c.grid;
    G.places.resize((size_t)D * nq);
    CYMG_HIP(hipMemcpyAsync(G.places.get(), places, (size_t)D * nq * 8, hipMemcpyHostToDevice, st));
    if (exclude) {
        G.excl.resize(nq);
        CYMG_HIP(hipMemcpyAsync(G.excl.get(), exclude, (size_t)nq * 4, hipMemcpyHostToDevice, st));
    }
    G.count.resize(nq + 1);
    G.offset.resize(nq + 1);
    G.flag.resize(1);
    G.half.resize(1);
    CYMG_HIP(hipMemsetAsync(G.half.get(), 0, 8, st));
    CYMG_HIP(hipMemsetAsync(G.flag.get(), 0, sizeof(int), st));
    const dim3 b(kGridThreads), gp((fib_pts + kGridThreads - 1) / kGridThreads), gs((nseg + kGridThreads - 1) / kGridThreads);
    if (fib_pts) {
        if (D == 3)
            hipLaunchKernelGGL(k_seg_half<3>, gp, b, 0, st, fib_pts, c.d_pfib.get(), c.d_idx.get(), c.d_np.get(),
                               c.d_pos.get(), G.half.get());
        else
            hipLaunchKernelGGL(k_seg_half<2>, gp, b, 0, st, fib_pts, c.d_pfib.get(), c.d_idx.get(), c.d_np.get(),
                               c.d_pos.get(), G.half.get());
        CYMG_HIP(hipGetLastError());
    }
    unsigned long long hb = 0;
    CYMG_HIP(hipMemcpyAsync(&hb, G.half.get(), 8, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipStreamSynchronize(st));
    double half = 0.0;
    std::memcpy(&half, &hb, 8);
    double cut_max = 0.0;
    for (uint32_t f = 0; f < c.nfib; ++f) cut_max = std::max(cut_max, c.h_cut[f]);
    double h = (std::sqrt(range * range + cut_max * cut_max) + half) * (1.0 + 1e-6);
    if (!(h > 0)) h = 1.0;
    // buckets: segments sorted by the hash of their midpoint's cell
    G.key.resize(nseg ? nseg : 1); G.key2.resize(nseg ? nseg : 1);
    G.val.resize(nseg ? nseg : 1); G.val2.resize(nseg ? nseg : 1);
    G.start.resize(kGridBuckets); G.end.resize(kGridBuckets);
    CYMG_HIP(hipMemsetAsync(G.end.get(), 0, (size_t)kGridBuckets * 4, st));
    if (nseg) {
        if (D == 3)
            hipLaunchKernelGGL(k_seg_keys<3>, gp, b, 0, st, fib_pts, c.d_pfib.get(), c.d_idx.get(), c.d_np.get(),
                               c.d_pos.get(), 1.0 / h, G.key.get(), G.val.get());
        else
            hipLaunchKernelGGL(k_seg_keys<2>, gp, b, 0, st, fib_pts, c.d_pfib.get(), c.d_idx.get(), c.d_np.get(),
                               c.d_pos.get(), 1.0 / h, G.key.get(), G.val.get());
        CYMG_HIP(hipGetLastError());
        size_t tb = 0;
        CYMG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, G.key.get(), G.key2.get(), G.val.get(), G.val2.get(),
                                                    (int)nseg, 0, kGridHashBits, st));
        G.temp.resize(std::max(tb, G.temp.n));
        CYMG_HIP(hipcub::DeviceRadixSort::SortPairs(G.temp.get(), tb, G.key.get(), G.key2.get(), G.val.get(),
                                                    G.val2.get(), (int)nseg, 0, kGridHashBits, st));
        hipLaunchKernelGGL(k_buckets, gs, b, 0, st, nseg, G.key2.get(), G.start.get(), G.end.get());
        CYMG_HIP(hipGetLastError());
    }
    GridQuery q;
    std::memset(&q, 0, sizeof(q));
    q.places = G.places.get();
    q.exclude = exclude ? G.excl.get() : nullptr;
    q.range2 = range * range;
    q.inclusive = inclusive;
    q.inv_h = 1.0 / h;
    q.sorted = G.val2.get();
    q.start = G.start.get();
    q.end = G.end.get();
    q.pfib = c.d_pfib.get();
    q.idx = c.d_idx.get();
    q.np = c.d_np.get();
    q.cut = c.d_cut.get();
    q.pos = c.d_pos.get();
    q.count = G.count.get();
    q.offset = G.offset.get();
    q.overflow = G.flag.get();
    const dim3 gq((nq + kGridWaves - 1) / kGridWaves);
    if (D == 3) hipLaunchKernelGGL((k_query_wave<3, false>), gq, b, 0, st, nq, q);
    else hipLaunchKernelGGL((k_query_wave<2, false>), gq, b, 0, st, nq, q);
    CYMG_HIP(hipGetLastError());
    CYMG_HIP(hipMemsetAsync(G.count.get() + nq, 0, 4, st));
    size_t tb = 0;
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, G.count.get(), G.offset.get(), (int)(nq + 1), st));
    G.temp.resize(std::max(tb, G.temp.n));
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(G.temp.get(), tb, G.count.get(), G.offset.get(), (int)(nq + 1), st));
    uint32_t n = 0;
    std::vector<uint32_t> cnt(nq);
    CYMG_HIP(hipMemcpyAsync(&n, G.offset.get() + nq, 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(cnt.data(), G.count.get(), (size_t)nq * 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipStreamSynchronize(st));
    // the fill pass ranks a place's segments in a list of kGridMaxKeep: refuse
    // before any output is written
    for (uint32_t k = 0; k < nq; ++k)
        if (cnt[k] > (uint32_t)kGridMaxKeep) throw Error(CYMG_ERR_UNSUPPORTED, "more than 64 segments near one place");
    std::copy(cnt.begin(), cnt.end(), counts);
    *total = n;
    if (n > capacity || !n) return;
    G.ofib.resize(n); G.opt.resize(n); G.oabs.resize(n); G.odis.resize(n);
    q.out_fib = G.ofib.get();
    q.out_pt = G.opt.get();
    q.out_abs = G.oabs.get();
    q.out_dis = G.odis.get();
    if (D == 3) hipLaunchKernelGGL((k_query_wave<3, true>), gq, b, 0, st, nq, q);
    else hipLaunchKernelGGL((k_query_wave<2, true>), gq, b, 0, st, nq, q);
    CYMG_HIP(hipGetLastError());
    int overflow = 0;
    CYMG_HIP(hipMemcpyAsync(&overflow, G.flag.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(fiber, G.ofib.get(), (size_t)n * 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(point, G.opt.get(), (size_t)n * 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(abscissa, G.oabs.get(), (size_t)n * 8, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(dist_sqr, G.odis.get(), (size_t)n * 8, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipStreamSynchronize(st));
    if (overflow) throw Error(CYMG_ERR_UNSUPPORTED, "more than 64 segments near one place");
}

}  // namespace cymg
