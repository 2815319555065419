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
#include <mutex>

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

// Device memory of one query: staging of the inputs, the bucket structure,
// the count/offset pass, the sort/scan work space and the output staging.
struct GridSpace {
    double* places;
    uint32_t *excl, *count, *offset, *key, *key2, *val, *val2, *start, *end;
    int* flag;
    unsigned long long* half;
    void* temp;
    size_t temp_bytes;
    uint32_t *ofib, *opt;
    double *oabs, *odis;
};

// Host side of cymg_nearby_segments.  Cell side h: a kept segment has
// |w - P1|^2 - abs^2 <= range^2 with 0 <= abs <= segmentation (or is within
// range of a fiber end), so its midpoint lies within sqrt(range^2 +
// segmentation^2) + half its length of the place; h is that bound for the
// longest segmentation and segment, enlarged by 1e-6 against rounding.
// `sp` provides the device memory: sp.inputs(G) before anything runs, then
// sp.temp(G, bytes) for the sort and scan, sp.outputs(G, n) for n kept
// segments (after the count pass, before any output is written).
template <class Space>
static void grid_query(Context& c, Space& sp, uint32_t nq, const double* places, const uint32_t* exclude,
                       double range, int inclusive, uint32_t* counts, uint32_t capacity, uint32_t* fiber,
                       uint32_t* point, double* abscissa, double* dist_sqr, uint32_t* total)
{
    hipStream_t st = c.stream;
    const int D = c.dim;
    const uint32_t fib_pts = c.fib_pts, nseg = c.nseg;
    if (!nq) {
        *total = 0;
        return;
    }
    GridSpace G;
    std::memset(&G, 0, sizeof(G));
    sp.inputs(G, nq, nseg, exclude != nullptr);
    CYMG_HIP(hipMemcpyAsync(G.places, places, (size_t)D * nq * 8, hipMemcpyHostToDevice, st));
    if (exclude) CYMG_HIP(hipMemcpyAsync(G.excl, exclude, (size_t)nq * 4, hipMemcpyHostToDevice, st));
    CYMG_HIP(hipMemsetAsync(G.half, 0, 8, st));
    CYMG_HIP(hipMemsetAsync(G.flag, 0, sizeof(int), st));
    const dim3 b(kGridThreads), gp((fib_pts + kGridThreads - 1) / kGridThreads), gs((nseg + kGridThreads - 1) / kGridThreads);
    if (fib_pts) {
        if (D == 3)
            hipLaunchKernelGGL(k_seg_half<3>, gp, b, 0, st, fib_pts, c.d_pfib.get(), c.d_idx.get(), c.d_np.get(),
                               c.d_pos.get(), G.half);
        else
            hipLaunchKernelGGL(k_seg_half<2>, gp, b, 0, st, fib_pts, c.d_pfib.get(), c.d_idx.get(), c.d_np.get(),
                               c.d_pos.get(), G.half);
        CYMG_HIP(hipGetLastError());
    }
    unsigned long long hb = 0;
    CYMG_HIP(hipMemcpyAsync(&hb, G.half, 8, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipStreamSynchronize(st));
    double half = 0.0;
    std::memcpy(&half, &hb, 8);
    double cut_max = 0.0;
    for (uint32_t f = 0; f < c.nfib; ++f) cut_max = std::max(cut_max, c.h_cut[f]);
    double h = (std::sqrt(range * range + cut_max * cut_max) + half) * (1.0 + 1e-6);
    if (!(h > 0)) h = 1.0;
    // buckets: segments sorted by the hash of their midpoint's cell
    CYMG_HIP(hipMemsetAsync(G.end, 0, (size_t)kGridBuckets * 4, st));
    if (nseg) {
        if (D == 3)
            hipLaunchKernelGGL(k_seg_keys<3>, gp, b, 0, st, fib_pts, c.d_pfib.get(), c.d_idx.get(), c.d_np.get(),
                               c.d_pos.get(), 1.0 / h, G.key, G.val);
        else
            hipLaunchKernelGGL(k_seg_keys<2>, gp, b, 0, st, fib_pts, c.d_pfib.get(), c.d_idx.get(), c.d_np.get(),
                               c.d_pos.get(), 1.0 / h, G.key, G.val);
        CYMG_HIP(hipGetLastError());
        size_t tb = 0;
        CYMG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, G.key, G.key2, G.val, G.val2, (int)nseg, 0,
                                                    kGridHashBits, st));
        sp.temp(G, tb);
        CYMG_HIP(hipcub::DeviceRadixSort::SortPairs(G.temp, tb, G.key, G.key2, G.val, G.val2, (int)nseg, 0,
                                                    kGridHashBits, st));
        hipLaunchKernelGGL(k_buckets, gs, b, 0, st, nseg, G.key2, G.start, G.end);
        CYMG_HIP(hipGetLastError());
    }
    GridQuery q;
    std::memset(&q, 0, sizeof(q));
    q.places = G.places;
    q.exclude = exclude ? G.excl : nullptr;
    q.range2 = range * range;
    q.inclusive = inclusive;
    q.inv_h = 1.0 / h;
    q.sorted = G.val2;
    q.start = G.start;
    q.end = G.end;
    q.pfib = c.d_pfib.get();
    q.idx = c.d_idx.get();
    q.np = c.d_np.get();
    q.cut = c.d_cut.get();
    q.pos = c.d_pos.get();
    q.count = G.count;
    q.offset = G.offset;
    q.overflow = G.flag;
    const dim3 gq((nq + kGridWaves - 1) / kGridWaves);
    if (D == 3) hipLaunchKernelGGL((k_query_wave<3, false>), gq, b, 0, st, nq, q);
    else hipLaunchKernelGGL((k_query_wave<2, false>), gq, b, 0, st, nq, q);
    CYMG_HIP(hipGetLastError());
    CYMG_HIP(hipMemsetAsync(G.count + nq, 0, 4, st));
    size_t tb = 0;
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, G.count, G.offset, (int)(nq + 1), st));
    sp.temp(G, tb);
    CYMG_HIP(hipcub::DeviceScan::ExclusiveSum(G.temp, tb, G.count, G.offset, (int)(nq + 1), st));
    uint32_t n = 0;
    std::vector<uint32_t> cnt(nq);
    CYMG_HIP(hipMemcpyAsync(&n, G.offset + nq, 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(cnt.data(), G.count, (size_t)nq * 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipStreamSynchronize(st));
    // the fill pass ranks a place's segments in a list of kGridMaxKeep: refuse
    // before any output is written
    for (uint32_t k = 0; k < nq; ++k)
        if (cnt[k] > (uint32_t)kGridMaxKeep) throw Error(CYMG_ERR_UNSUPPORTED, "more than 64 segments near one place");
    if (n <= capacity && n) sp.outputs(G, n);
    std::copy(cnt.begin(), cnt.end(), counts);
    *total = n;
    if (n > capacity || !n) return;
    q.out_fib = G.ofib;
    q.out_pt = G.opt;
    q.out_abs = G.oabs;
    q.out_dis = G.odis;
    if (D == 3) hipLaunchKernelGGL((k_query_wave<3, true>), gq, b, 0, st, nq, q);
    else hipLaunchKernelGGL((k_query_wave<2, true>), gq, b, 0, st, nq, q);
    CYMG_HIP(hipGetLastError());
    int overflow = 0;
    CYMG_HIP(hipMemcpyAsync(&overflow, G.flag, sizeof(int), hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(fiber, G.ofib, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(point, G.opt, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(abscissa, G.oabs, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipMemcpyAsync(dist_sqr, G.odis, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    CYMG_HIP(hipStreamSynchronize(st));
    if (overflow) throw Error(CYMG_ERR_UNSUPPORTED, "more than 64 segments near one place");
}

// the context's grow-only buffers (cymg_nearby_segments)
struct ContextSpace {
    GridBufs& B;
    int D;
    void inputs(GridSpace& G, uint32_t nq, uint32_t nseg, bool excl)
    {
        B.places.resize((size_t)D * nq);
        if (excl) B.excl.resize(nq);
        B.count.resize(nq + 1);
        B.offset.resize(nq + 1);
        B.flag.resize(1);
        B.half.resize(1);
        const size_t ns = nseg ? nseg : 1;
        B.key.resize(ns); B.key2.resize(ns); B.val.resize(ns); B.val2.resize(ns);
        B.start.resize(kGridBuckets); B.end.resize(kGridBuckets);
        G.places = B.places.get();
        G.excl = B.excl.get();
        G.count = B.count.get();
        G.offset = B.offset.get();
        G.flag = B.flag.get();
        G.half = B.half.get();
        G.key = B.key.get(); G.key2 = B.key2.get(); G.val = B.val.get(); G.val2 = B.val2.get();
        G.start = B.start.get(); G.end = B.end.get();
    }
    void temp(GridSpace& G, size_t bytes)
    {
        B.temp.resize(std::max(bytes, B.temp.n));
        G.temp = B.temp.get();
        G.temp_bytes = B.temp.n;
    }
    void outputs(GridSpace& G, uint32_t n)
    {
        B.ofib.resize(n); B.opt.resize(n); B.oabs.resize(n); B.odis.resize(n);
        G.ofib = B.ofib.get(); G.opt = B.opt.get(); G.oabs = B.oabs.get(); G.odis = B.odis.get();
    }
};

void nearby_segments(Context& c, uint32_t nq, const double* places, const uint32_t* exclude, double range,
                     int inclusive, uint32_t* counts, uint32_t capacity, uint32_t* fiber, uint32_t* point,
                     double* abscissa, double* dist_sqr, uint32_t* total)
{
    ContextSpace sp{ c.grid, c.dim };
    grid_query(c, sp, nq, places, exclude, range, inclusive, counts, capacity, fiber, point, abscissa, dist_sqr,
               total);
}

// ---------------------------------------------------------------------------
// The same query without any allocation (cymg_nearby_segments_noalloc): all
// of its device memory is carved out of one file-scope array per GPU,
// `scratch`, whose device address each GPU's first context looks up once
// (grid_scratch_init, from cymg_create).  Host code never uses the array's
// name as an address; it uses the looked-up pointer.  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU (two streams, two
// threads) race.  What must start at zero (the bucket ends, the half-length
// maximum, the overflow flag, the sentinel count) is cleared by every call.
constexpr size_t kScratchAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }
constexpr size_t kNaPlaces = CYMG_NOALLOC_MAX_PLACES, kNaSegs = CYMG_NOALLOC_MAX_SEGMENTS,
                 kNaKept = CYMG_NOALLOC_MAX_KEPT;
constexpr size_t kOffPlaces = 0;
constexpr size_t kOffExcl = kOffPlaces + align_up(3 * 8 * kNaPlaces);
constexpr size_t kOffCount = kOffExcl + align_up(4 * kNaPlaces);
constexpr size_t kOffOffset = kOffCount + align_up(4 * (kNaPlaces + 1));
constexpr size_t kOffFlag = kOffOffset + align_up(4 * (kNaPlaces + 1));
constexpr size_t kOffHalf = kOffFlag + kScratchAlign;
constexpr size_t kOffKey = kOffHalf + kScratchAlign;
constexpr size_t kOffKey2 = kOffKey + align_up(4 * kNaSegs);
constexpr size_t kOffVal = kOffKey2 + align_up(4 * kNaSegs);
constexpr size_t kOffVal2 = kOffVal + align_up(4 * kNaSegs);
constexpr size_t kOffStart = kOffVal2 + align_up(4 * kNaSegs);
constexpr size_t kOffEnd = kOffStart + align_up(4 * (size_t)kGridBuckets);
constexpr size_t kOffTemp = kOffEnd + align_up(4 * (size_t)kGridBuckets);
constexpr size_t kTempBytes = 4u << 20;
constexpr size_t kOffFib = kOffTemp + kTempBytes;
constexpr size_t kOffPt = kOffFib + align_up(4 * kNaKept);
constexpr size_t kOffAbs = kOffPt + align_up(4 * kNaKept);
constexpr size_t kOffDis = kOffAbs + align_up(8 * kNaKept);
constexpr size_t kScratchBytes = kOffDis + align_up(8 * kNaKept);
static_assert(kScratchBytes <= CYMG_NOALLOC_SCRATCH_BYTES, "the no-allocation query's regions exceed its array");

static __device__ __align__(256) unsigned char scratch[CYMG_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

// once per device, with that device current: the device address of its copy
// of `scratch`
void grid_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) throw Error(CYMG_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw Error(CYMG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
}

struct ScratchSpace {
    unsigned char* base;
    void inputs(GridSpace& G, uint32_t, uint32_t, bool)
    {
        G.places = reinterpret_cast<double*>(base + kOffPlaces);
        G.excl = reinterpret_cast<uint32_t*>(base + kOffExcl);
        G.count = reinterpret_cast<uint32_t*>(base + kOffCount);
        G.offset = reinterpret_cast<uint32_t*>(base + kOffOffset);
        G.flag = reinterpret_cast<int*>(base + kOffFlag);
        G.half = reinterpret_cast<unsigned long long*>(base + kOffHalf);
        G.key = reinterpret_cast<uint32_t*>(base + kOffKey);
        G.key2 = reinterpret_cast<uint32_t*>(base + kOffKey2);
        G.val = reinterpret_cast<uint32_t*>(base + kOffVal);
        G.val2 = reinterpret_cast<uint32_t*>(base + kOffVal2);
        G.start = reinterpret_cast<uint32_t*>(base + kOffStart);
        G.end = reinterpret_cast<uint32_t*>(base + kOffEnd);
    }
    void temp(GridSpace& G, size_t bytes)
    {
        if (bytes > kTempBytes) throw Error(CYMG_ERR_UNSUPPORTED, "sort work space beyond the no-allocation array");
        G.temp = base + kOffTemp;
        G.temp_bytes = kTempBytes;
    }
    void outputs(GridSpace& G, uint32_t n)
    {
        if (n > kNaKept)
            throw Error(CYMG_ERR_UNSUPPORTED, "more than CYMG_NOALLOC_MAX_KEPT segments kept in one call");
        G.ofib = reinterpret_cast<uint32_t*>(base + kOffFib);
        G.opt = reinterpret_cast<uint32_t*>(base + kOffPt);
        G.oabs = reinterpret_cast<double*>(base + kOffAbs);
        G.odis = reinterpret_cast<double*>(base + kOffDis);
    }
};

void nearby_segments_noalloc(Context& c, uint32_t nq, const double* places, const uint32_t* exclude, double range,
                             int inclusive, uint32_t* counts, uint32_t capacity, uint32_t* fiber, uint32_t* point,
                             double* abscissa, double* dist_sqr, uint32_t* total)
{
    if (nq > kNaPlaces) throw Error(CYMG_ERR_UNSUPPORTED, "more than CYMG_NOALLOC_MAX_PLACES places");
    if (c.nseg > kNaSegs) throw Error(CYMG_ERR_UNSUPPORTED, "more than CYMG_NOALLOC_MAX_SEGMENTS fiber segments");
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        base = (c.device >= 0 && c.device < kMaxDevices) ? g_scratch_ptr[c.device] : nullptr;
    }
    if (!base) throw Error(CYMG_ERR_INVALID, "no-allocation array not set up on this device");
    ScratchSpace sp{ base };
    grid_query(c, sp, nq, places, exclude, range, inclusive, counts, capacity, fiber, point, abscissa, dist_sqr,
               total);
}

}  // namespace cymg
