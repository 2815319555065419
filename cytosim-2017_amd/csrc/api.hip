// C ABI of the MI355X Meca library (include/cymeca.h): context, registration,
// host-side assembly recording in call order, and the solve orchestration with
// the reference's rescue chain.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "context.h"
#include "solid_ops.h"
#include "tile_ops.h"

using namespace cymg;

struct cymg_context {
    Context c;
};

namespace {

template <typename F>
int guarded(cymg_context* ctx, F&& fn)
{
    if (!ctx) {
        set_thread_error("null context");
        return CYMG_ERR_INVALID;
    }
    try {
        CYMG_HIP(hipSetDevice(ctx->c.device));
        fn(ctx->c);
        ctx->c.err.clear();
        return CYMG_OK;
    } catch (const Error& e) {
        ctx->c.err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->c.err = e.what();
        return CYMG_ERR_INVALID;
    }
}

inline uint64_t lower_key_h(uint32_t i, uint32_t j)
{
    const uint32_t r = std::max(i, j), c = std::min(i, j);
    return ((uint64_t)c << 32) | r;
}

void addB(Context& c, uint32_t i, uint32_t j, double delta)
{
    c.hB.key.push_back(lower_key_h(i, j));
    c.hB.val.push_back(delta);
}

void addC(Context& c, uint32_t i, uint32_t j, double delta)
{
    c.hC.key.push_back(lower_key_h(i, j));
    c.hC.val.push_back(delta);
}

void addBas(Context& c, uint32_t i, double delta)
{
    c.hBas.key.push_back(i);
    c.hBas.val.push_back(delta);
}

void need_prepared(const Context& c)
{
    if (!c.prepared) throw Error(CYMG_ERR_INVALID, "call cymg_prepare before assembling interactions");
}

void check_point(const Context& c, uint32_t a)
{
    if (a >= c.npts) throw Error(CYMG_ERR_INVALID, "point index out of range");
}

int mecable_of(const Context& c, uint32_t a)
{
    if (a >= c.fib_pts && c.nsol) {
        auto it = std::upper_bound(c.h_sol_idx.begin(), c.h_sol_idx.end(), a);
        return (int)c.nfib + (int)(it - c.h_sol_idx.begin()) - 1;
    }
    auto it = std::upper_bound(c.h_idx.begin(), c.h_idx.end(), a);
    return (int)(it - c.h_idx.begin()) - 1;
}

int system_of_fiber(const Context& c, uint32_t f)
{
    if (c.h_sys_begin.size() < 2) return 0;
    auto it = std::upper_bound(c.h_sys_begin.begin(), c.h_sys_begin.end(), f);
    return (int)(it - c.h_sys_begin.begin()) - 1;
}

// interactions may not couple two independent systems
void same_system(const Context& c, std::initializer_list<uint32_t> pts)
{
    if (c.h_sys_begin.size() < 3) return;
    int s0 = -1;
    for (uint32_t p : pts) {
        const int s = system_of_fiber(c, (uint32_t)mecable_of(c, p));
        if (s0 < 0) s0 = s;
        else if (s != s0) throw Error(CYMG_ERR_INVALID, "interaction couples two independent systems");
    }
}

// Tiles of consecutive whole fibers of one system, at most `target` points
// (a longer fiber gets a tile of its own) and kTileMaxFibers fibers.  The
// default target is the largest tile whose unknowns get two threads each in
// the iteration launches (256 / DIM points): few enough tiles that one round
// of workgroups covers the 1000-fiber system on 256 compute units.
void build_tiles(Context& c)
{
    const std::vector<uint32_t>& sb = c.h_sys_begin_used;
    // one tile per Solid after the fiber tiles (a layout with Solids is one
    // system), when every Solid fits a workgroup of the iteration launches
    c.solid_tiles = c.nsol > 0;
    for (uint32_t k = 0; k < c.nsol; ++k)
        if (c.dim * c.h_sol_np[k] > 1024u) c.solid_tiles = false;
    // the tile limits of the iteration's workgroup: tiles of at most
    // threads/2 unknowns give every unknown two threads (the default), larger
    // ones one thread; a fiber longer than the target gets a tile of its own
    const int threads = (c.tile_threads == 256 && (c.nsol == 0 || c.solid_tiles) &&
                         c.largest <= (uint32_t)tile_max_points(c, 256)) ? 256 : 512;
    const uint32_t maxp = (uint32_t)tile_max_points(c, threads);
    uint32_t target = c.tile_points;
    if (!target) target = std::min<uint32_t>((uint32_t)threads / 2 / (uint32_t)c.dim, kMaxFiberPoints);
    target = std::min(target, maxp);
    c.h_tile_fib.clear();
    c.h_tile_sys.clear();
    c.h_sys_tile.assign(c.nsys + 1, 0);
    for (uint32_t s = 0; s < c.nsys; ++s) {
        c.h_sys_tile[s] = (uint32_t)c.h_tile_fib.size();
        uint32_t pts = 0, nf = 0;
        for (uint32_t f = sb[s]; f < sb[s + 1]; ++f) {
            if (nf && (pts + c.h_np[f] > target || nf == (uint32_t)kTileMaxFibers)) nf = 0;
            if (!nf) {
                c.h_tile_fib.push_back(f);
                c.h_tile_sys.push_back(s);
                pts = 0;
            }
            pts += c.h_np[f];
            ++nf;
        }
    }
    const uint32_t nft = (uint32_t)c.h_tile_fib.size();
    c.ntiles = nft + (c.solid_tiles ? c.nsol : 0);
    c.h_sys_tile[c.nsys] = c.ntiles;
    c.h_tile_fib.push_back(c.nfib);
    c.h_fslot.assign(c.nfib ? c.nfib : 1, 0);
    std::vector<TileDesc> desc(c.ntiles);
    for (uint32_t t = 0; t < c.ntiles; ++t) {
        TileDesc& d = desc[t];
        if (t >= nft) {
            const uint32_t k = t - nft;
            d.fb = c.nfib + k;       // the Solid's Mecable
            d.nft = 0;               // no fiber: a Solid tile
            d.P0 = c.h_sol_idx[k];
            d.tp = c.h_sol_np[k];
            d.sys = 0;
            d.tb = c.h_sys_tile[0];
            d.te = c.h_sys_tile[1];
            d.S0 = 0;
            d.ws = 0;
            continue;
        }
        const uint32_t fb = c.h_tile_fib[t], fe = c.h_tile_fib[t + 1];
        d.fb = fb;
        d.nft = fe - fb;
        d.P0 = c.h_idx[fb];
        d.tp = c.h_idx[fe - 1] + c.h_np[fe - 1] - d.P0;
        d.sys = c.h_tile_sys[t];
        d.tb = c.h_sys_tile[d.sys];
        d.te = c.h_sys_tile[d.sys + 1];
        d.S0 = c.h_seg[fb];
        // segment lanes: the tile's fibers in order, a fiber moved to the next
        // wavefront when its segments would cross one, so that each wavefront
        // solves whole tridiagonal systems; not packed when a fiber has more
        // segments than a wavefront has lanes or the lanes run out
        uint32_t lane = 0;
        bool packed = true;
        for (uint32_t f = fb; f < fe; ++f) {
            const uint32_t S = c.h_np[f] - 1;
            if (S > (uint32_t)kWave) packed = false;
            if (lane % kWave + S > (uint32_t)kWave) lane = (lane / kWave + 1) * kWave;
            c.h_fslot[f] = lane;
            lane += S;
        }
        d.ws = (packed && lane <= (uint32_t)threads) ? lane : 0;
    }
    c.d_tile_desc.resize(c.ntiles ? c.ntiles : 1);
    c.d_tpart.resize((size_t)P_SLOTS * (c.ntiles ? c.ntiles : 1));
    hipStream_t st = c.stream;
    c.d_fslot.resize(c.h_fslot.size());
    CYMG_HIP(hipMemcpyAsync(c.d_fslot.get(), c.h_fslot.data(), c.h_fslot.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, st));
    if (c.ntiles)
        CYMG_HIP(hipMemcpyAsync(c.d_tile_desc.get(), desc.data(), c.ntiles * sizeof(TileDesc), hipMemcpyHostToDevice,
                                st));
    CYMG_HIP(hipStreamSynchronize(st));   // pageable source
    c.last_iters[0] = c.last_iters[1] = c.last_iters[2] = 0;
}

// Point map from the previous registration to the current one: each Mecable
// takes the points of the Mecable with the same identity and point count, or,
// when neither registration gave identities, of the one at the same position
// if every point count is unchanged.  Returns whether any point is mapped;
// `all` tells whether every point is.
bool remap_plan(const Context& c, std::vector<uint32_t>& map, bool& all)
{
    all = false;
    if (!c.prev_layout) return false;
    const uint32_t M = c.nmec();
    std::vector<uint32_t> np(M), idx(M);
    for (uint32_t f = 0; f < c.nfib; ++f) { np[f] = c.h_np[f]; idx[f] = c.h_idx[f]; }
    for (uint32_t k = 0; k < c.nsol; ++k) { np[c.nfib + k] = c.h_sol_np[k]; idx[c.nfib + k] = c.h_sol_idx[k]; }
    const std::vector<uint32_t>& onp = c.h_prev_np;
    std::vector<uint32_t> oidx(onp.size() + 1, 0);
    for (size_t m = 0; m < onp.size(); ++m) oidx[m + 1] = oidx[m] + onp[m];
    std::vector<uint32_t> src(M, kInvalid32);   // previous Mecable of each Mecable
    if (!c.h_ids.empty() && !c.h_prev_ids.empty()) {
        std::unordered_map<uint32_t, uint32_t> old;
        old.reserve(c.h_prev_ids.size() * 2);
        for (uint32_t m = 0; m < (uint32_t)c.h_prev_ids.size(); ++m) old.emplace(c.h_prev_ids[m], m);
        for (uint32_t m = 0; m < M; ++m) {
            auto it = old.find(c.h_ids[m]);
            if (it != old.end() && onp[it->second] == np[m]) src[m] = it->second;
        }
    } else if (c.h_ids.empty() && c.h_prev_ids.empty() && onp == np) {
        for (uint32_t m = 0; m < M; ++m) src[m] = m;
    }
    map.assign(c.npts, kInvalid32);
    bool any = false;
    all = true;
    for (uint32_t m = 0; m < M; ++m) {
        if (src[m] == kInvalid32) {
            all = false;
            continue;
        }
        any = true;
        for (uint32_t p = 0; p < np[m]; ++p) map[idx[m] + p] = oidx[src[m]] + p;
    }
    return any;
}

void upload_layout(Context& c, bool need_points)
{
    ++c.layout_gen;
    const uint32_t F = c.nfib;
    c.d_drag.resize(3 * (size_t)(F ? F : 1));
    if (F) CYMG_HIP(hipMemcpyAsync(c.d_drag.get(), c.h_drag.data(), 3 * (size_t)F * 8, hipMemcpyHostToDevice,
                                   c.stream));
    c.d_np.resize(F); c.d_idx.resize(F); c.d_seg.resize(F); c.d_cut.resize(F); c.d_rig.resize(F);
    c.d_mob.resize(F); c.d_rfrig.resize(F); c.d_blkoff.resize(F + 1); c.d_fib_sys.resize(F);
    std::vector<uint32_t> sb = c.h_sys_begin;
    if (sb.size() < 2 || sb.back() != F) sb = { 0, F };
    c.nsys = (uint32_t)sb.size() - 1;
    c.d_sys_begin.resize(sb.size());
    std::vector<uint32_t> sysv(F, 0);
    for (uint32_t s = 0; s < c.nsys; ++s)
        for (uint32_t f = sb[s]; f < sb[s + 1]; ++f) sysv[f] = s;
    hipStream_t st = c.stream;
    CYMG_HIP(hipMemcpyAsync(c.d_np.get(), c.h_np.data(), F * 4, hipMemcpyHostToDevice, st));
    CYMG_HIP(hipMemcpyAsync(c.d_idx.get(), c.h_idx.data(), F * 4, hipMemcpyHostToDevice, st));
    CYMG_HIP(hipMemcpyAsync(c.d_seg.get(), c.h_seg.data(), F * 4, hipMemcpyHostToDevice, st));
    CYMG_HIP(hipMemcpyAsync(c.d_cut.get(), c.h_cut.data(), F * 8, hipMemcpyHostToDevice, st));
    CYMG_HIP(hipMemcpyAsync(c.d_rig.get(), c.h_rig.data(), F * 8, hipMemcpyHostToDevice, st));
    CYMG_HIP(hipMemcpyAsync(c.d_blkoff.get(), c.h_blkoff.data(), (F + 1) * 8, hipMemcpyHostToDevice, st));
    CYMG_HIP(hipMemcpyAsync(c.d_fib_sys.get(), sysv.data(), F * 4, hipMemcpyHostToDevice, st));
    CYMG_HIP(hipMemcpyAsync(c.d_sys_begin.get(), sb.data(), sb.size() * 4, hipMemcpyHostToDevice, st));
    c.h_sys_begin_used = sb;
    if (c.nsol) {
        // one system: its per-Mecable partial sums cover the Solids after the fibers
        const std::vector<uint32_t> sbm = { 0, c.nmec() };
        CYMG_HIP(hipMemcpyAsync(c.d_sys_begin.get(), sbm.data(), 8, hipMemcpyHostToDevice, st));
        const uint32_t K = c.nsol, SP = c.npts - c.fib_pts;
        c.d_sol_np.resize(K); c.d_sol_idx.resize(K); c.d_sol_shape_sqr.resize(K); c.d_sol_timer.resize(K);
        c.d_sol_rad.resize(SP); c.d_sol_shape.resize((size_t)c.dim * SP);
        CYMG_HIP(hipMemcpyAsync(c.d_sol_np.get(), c.h_sol_np.data(), K * 4, hipMemcpyHostToDevice, st));
        CYMG_HIP(hipMemcpyAsync(c.d_sol_idx.get(), c.h_sol_idx.data(), K * 4, hipMemcpyHostToDevice, st));
        CYMG_HIP(hipMemcpyAsync(c.d_sol_shape_sqr.get(), c.h_sol_shape_sqr.data(), K * 8, hipMemcpyHostToDevice, st));
        CYMG_HIP(hipMemcpyAsync(c.d_sol_timer.get(), c.h_sol_timer.data(), K * 4, hipMemcpyHostToDevice, st));
        CYMG_HIP(hipMemcpyAsync(c.d_sol_rad.get(), c.h_sol_rad.data(), SP * 8, hipMemcpyHostToDevice, st));
        CYMG_HIP(hipMemcpyAsync(c.d_sol_shape.get(), c.h_sol_shape.data(), (size_t)c.dim * SP * 8,
                                hipMemcpyHostToDevice, st));
    }
    std::vector<uint32_t> pf(c.npts);
    for (uint32_t f = 0; f < F; ++f)
        for (uint32_t p = 0; p < c.h_np[f]; ++p) pf[c.h_idx[f] + p] = f;
    c.d_pfib.resize(c.npts ? c.npts : 1);
    if (c.npts) CYMG_HIP(hipMemcpyAsync(c.d_pfib.get(), pf.data(), c.npts * 4, hipMemcpyHostToDevice, st));
    const size_t n = c.n();
    // the previous registration's points and solution, carried over per Mecable
    std::vector<uint32_t> map;
    bool all = false;
    const bool any = remap_plan(c, map, all);
    if (any) {
        c.d_remap.resize(c.npts);
        CYMG_HIP(hipMemcpyAsync(c.d_remap.get(), map.data(), (size_t)c.npts * 4, hipMemcpyHostToDevice, st));
        for (DevBuf<double>* v : { &c.d_pos, &c.d_prev_sol }) {
            if (v == &c.d_pos && !(need_points && c.have_solution)) continue;
            c.d_old.swap(*v);
            v->resize(n ? n : 1);
            launch_gather_points(c, c.d_remap.get(), c.d_old.get(), v->get(), c.npts);
        }
    } else {
        c.d_prev_sol.resize(n ? n : 1);
        CYMG_HIP(hipMemsetAsync(c.d_prev_sol.get(), 0, (n ? n : 1) * 8, st));
    }
    c.pos_complete = any && all && c.have_solution;
    c.h_prev_ids = c.h_ids;
    c.h_prev_np.assign(c.h_np.begin(), c.h_np.end());
    c.h_prev_np.insert(c.h_prev_np.end(), c.h_sol_np.begin(), c.h_sol_np.end());
    c.prev_layout = true;
    for (DevBuf<double>* v : { &c.d_pts, &c.d_pos, &c.d_sol, &c.d_bas, &c.d_rhs, &c.d_for, &c.d_tmp, &c.d_prev_sol })
        v->resize(n ? n : 1);
    const size_t S = c.nseg ? c.nseg : 1;
    for (DevBuf<double>* v : { &c.d_diff }) v->resize(S * c.dim);
    for (DevBuf<double>* v : { &c.d_jd, &c.d_je, &c.d_jd0, &c.d_je0, &c.d_lag, &c.d_pdiff }) v->resize(S);
    c.d_pcr.resize(S * (2 * kPcrLevels + 1));
    CYMG_HIP(hipMemsetAsync(c.d_lag.get(), 0, S * 8, st));
    c.d_part.resize(P_SLOTS * (size_t)(c.nmec() ? c.nmec() : 1));
    c.d_scal.resize(2 * c.nsys);
    CYMG_HIP(hipMemsetAsync(c.d_scal.get(), 0, 2 * c.nsys * sizeof(SysScalars), st));
    CYMG_HIP(hipStreamSynchronize(st));   // the sources above are pageable host arrays
    build_tiles(c);
    c.layout_dirty = false;
}

struct Timer {
    Context& c;
    int slot;
    Timer(Context& cc, int s) : c(cc), slot(s) {
        if (c.timing) CYMG_HIP(hipEventRecord(c.ev[2 * slot], c.stream));
    }
    ~Timer() {
        if (c.timing) (void)hipEventRecord(c.ev[2 * slot + 1], c.stream);
    }
};

// 0 for a phase not (yet) recorded; the failed query's error is cleared so
// that it does not surface in a later error check
float elapsed(Context& c, int slot)
{
    float ms = 0;
    if (hipEventElapsedTime(&ms, c.ev[2 * slot], c.ev[2 * slot + 1]) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ms;
}

void do_setup(Context& c, const double* gauss, int on_device)
{
    if (!c.prepared) throw Error(CYMG_ERR_INVALID, "cymg_prepare must precede the solve");
    const size_t n = c.n();
    {
        Timer t(c, 0);
        build_streams(c);
    }
    const double* g = nullptr;
    if (gauss) {
        if (on_device) g = gauss;
        else {
            c.d_gauss.resize(n);
            CYMG_HIP(hipMemcpyAsync(c.d_gauss.get(), gauss, n * 8, hipMemcpyHostToDevice, c.stream));
            g = c.d_gauss.get();
        }
    }
    {
        Timer t(c, 1);
        launch_setup(c, g);
    }
    c.setup_done = true;
    c.pc_ready = false;
    c.solved = false;
}

// d_status: [0] fibers whose projection failed, [1] Solids that could not be
// prepared, [2] Solids that collapsed in the last getPoints.  prepare queues a
// copy into pinned memory (h_status); the solve reads it after its first
// synchronisation with the stream, so no extra round trip stalls the device.
void check_status(Context& c)
{
    const int* st = c.h_status;
    if (st[0]) throw Error(CYMG_ERR_PROJECTION, "could not build Fiber projection");
    // InvalidParameter / ABORT_NOW of the reference (solid.cc:690-703, 843-846)
    if (st[1]) throw Error(CYMG_ERR_INVALID, "zero drag in Solid, or failed to factorize Solid momentum matrix");
}

void do_iterate(Context& c, const cymg_solve_params* p, cymg_solve_report* rep)
{
    if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede cymg_iterate");
    cymg_solve_report r{};
    const int maxit = p->max_iter > 0 ? p->max_iter : (int)c.n();
    const int pc = p->precondition;
    if (pc < 0 || pc > 2) throw Error(CYMG_ERR_INVALID, "precondition must be 0, 1 or 2");
    if (p->guess < CYMG_GUESS_ZERO || p->guess > CYMG_GUESS_LEARNED) throw Error(CYMG_ERR_INVALID, "unknown guess");
    const double* guess = nullptr;
    int guess_used = CYMG_GUESS_ZERO;
    if (p->guess == CYMG_GUESS_PREVIOUS) {
        guess = c.d_prev_sol.get();
        guess_used = CYMG_GUESS_PREVIOUS;
    }
    if (pc) {
        Timer t(c, 2);
        launch_precond_build(c, pc);
    }
    // the learned guess and the recording of training samples need the
    // features of this step and M^-1 b
    const bool learned = p->guess == CYMG_GUESS_LEARNED && c.mlp.ready;
    c.mlp.step_ready = false;
    if (learned || c.mlp.recording) mlp_prepare_step(c, pc);
    if (learned) {
        guess = mlp_guess(c);
        guess_used = CYMG_GUESS_LEARNED;
    }
    {
        Timer t(c, 3);
        krylov_solve(c, pc, guess, maxit, p->tolerance, p->tol_mode, &r);
        check_status(c);   // the solve has synchronised past the status copy of prepare
        if (learned && mlp_guess_dropped(c)) guess_used = CYMG_GUESS_ZERO;
        if (r.converged && c.mlp.recording) mlp_record(c);
        int total = r.iterations;
        if (!r.converged && two_launch(c)) {
            // The two-launch iteration carries M^-1 r, M^-1 v and M^-1 t by
            // recurrences, whose rounding can hold the residual just above a
            // tolerance near the rounding floor; the same BCGSP with M^-1 applied
            // directly (four launches) continues from the iterate.
            CYMG_HIP(hipMemcpyAsync(c.d_tmp.get(), c.d_sol.get(), (size_t)c.n() * 8, hipMemcpyDeviceToDevice,
                                    c.stream));
            c.krylov_fused = false;
            cymg_solve_report r1{};
            try {
                krylov_solve(c, pc, c.d_tmp.get(), maxit, p->tolerance, p->tol_mode, &r1);
            } catch (...) {
                c.krylov_fused = true;
                throw;
            }
            c.krylov_fused = true;
            total += r1.iterations;
            r1.iterations += r.iterations;
            r = r1;
        }
        if (!r.converged && p->fallback) {
            // meca.cc:749-799: retry with vRHS as the seed, then switch method
            cymg_solve_report r2{};
            krylov_solve(c, pc, c.d_rhs.get(), maxit, p->tolerance, p->tol_mode, &r2);
            total += r2.iterations;
            if (r2.converged) {
                r = r2;
                r.fallback = 1;
            } else {
                cymg_solve_report r3{};
                int other = CYMG_PRECOND_NONE;
                if (!pc) {
                    launch_precond_build(c, CYMG_PRECOND_FP64);
                    other = CYMG_PRECOND_FP64;
                }
                krylov_solve(c, other, nullptr, maxit, p->tolerance, p->tol_mode, &r3);
                total += r3.iterations;
                r = r3;
                r.fallback = r3.converged ? 2 : 3;
            }
        }
        r.total_iterations = total;
    }
    r.guess_used = guess_used;
    if (rep) *rep = r;
    if (r.fallback == 3) throw Error(CYMG_ERR_CONVERGENCE, "convergence failure in solver");
    c.solved = true;
}

void do_apply(Context& c)
{
    Timer t(c, 4);
    launch_apply(c);
    CYMG_HIP(hipMemcpyAsync(c.d_prev_sol.get(), c.d_sol.get(), (size_t)c.n() * 8, hipMemcpyDeviceToDevice,
                            c.stream));
    c.have_solution = true;
}

void collect_timings(Context& c)
{
    if (!c.timing) return;
    CYMG_HIP(hipStreamSynchronize(c.stream));
    c.tm.assemble_ms = elapsed(c, 0);
    c.tm.setup_ms = elapsed(c, 1);
    c.tm.precond_ms = elapsed(c, 2);
    c.tm.iterate_ms = elapsed(c, 3);
    c.tm.apply_ms = elapsed(c, 4);
}

}  // namespace

extern "C" {

int cymg_create(int dim, int device, cymg_context** out)
{
    if (!out) return CYMG_ERR_INVALID;
    *out = nullptr;
    if (dim != 2 && dim != 3) return CYMG_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return CYMG_ERR_HIP;
    cymg_context* ctx = new cymg_context();
    ctx->c.dim = dim;
    ctx->c.device = device;
    int rc = guarded(ctx, [&](Context& c) {
        CYMG_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        CYMG_HIP(hipHostMalloc((void**)&c.h_scal, sizeof(SysScalars) * 4096, hipHostMallocDefault));
        CYMG_HIP(hipHostMalloc((void**)&c.h_snap, sizeof(SysScalars) * 2 * 4096, hipHostMallocDefault));
        CYMG_HIP(hipHostMalloc((void**)&c.h_status, 4 * sizeof(int), hipHostMallocDefault));
        for (int i = 0; i < 10; ++i) CYMG_HIP(hipEventCreate(&c.ev[i]));
        for (int i = 0; i < 2; ++i) CYMG_HIP(hipEventCreateWithFlags(&c.chunk_ev[i], hipEventDisableTiming));
        grid_scratch_init(c.device);   // the no-allocation query's array on this device
    });
    if (rc) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return CYMG_OK;
}

int cymg_destroy(cymg_context* ctx)
{
    if (!ctx) {
        set_thread_error("null context");
        return CYMG_ERR_INVALID;
    }
    Context& c = ctx->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (int i = 0; i < 10; ++i)
        if (c.ev[i]) (void)hipEventDestroy(c.ev[i]);
    if (c.h_scal) (void)hipHostFree(c.h_scal);
    if (c.h_snap) (void)hipHostFree(c.h_snap);
    if (c.h_status) (void)hipHostFree(c.h_status);
    if (c.h_st_tot) (void)hipHostFree(c.h_st_tot);
    if (c.mlp.h_gflag) (void)hipHostFree(c.mlp.h_gflag);
    for (int i = 0; i < 2; ++i)
        if (c.chunk_ev[i]) (void)hipEventDestroy(c.chunk_ev[i]);
    for (auto& row : c.iter_graph)
        for (hipGraphExec_t& g : row)
            if (g) (void)hipGraphExecDestroy(g);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete ctx;
    return CYMG_OK;
}

const char* cymg_last_error(const cymg_context* ctx)
{
    return ctx ? ctx->c.err.c_str() : thread_error();
}

int cymg_set_params(cymg_context* ctx, const cymg_sim_params* p)
{
    return guarded(ctx, [&](Context& c) {
        if (!p || !(p->time_step > 0) || !(p->viscosity > 0) || !(p->hydro_radius0 > 0) || !(p->hydro_radius1 > 0) ||
            p->kT < 0)
            throw Error(CYMG_ERR_INVALID, "invalid simulation parameters");
        c.prm = *p;
    });
}

int cymg_clear(cymg_context* ctx)
{
    return guarded(ctx, [&](Context& c) {
        c.h_np.clear(); c.h_idx.clear(); c.h_seg.clear(); c.h_cut.clear(); c.h_rig.clear(); c.h_drag.clear();
        c.h_ids.clear();
        c.h_blkoff.assign(1, 0);
        c.nfib = c.npts = c.nseg = c.largest = 0;
        c.nsol = c.fib_pts = 0;
        c.h_sol_np.clear(); c.h_sol_idx.clear(); c.h_sol_rad.clear(); c.h_sol_shape.clear();
        c.h_sol_shape_sqr.clear(); c.h_sol_timer.clear();
        c.blk_total = 0;
        c.h_sys_begin.clear();
        c.layout_dirty = true;
        c.prepared = c.setup_done = c.pc_ready = c.solved = false;
    });
}

static void add_fiber(Context& c, uint32_t np, double segmentation, double rigidity, const double drag[3],
                      uint32_t* mat_index)
{
        if (np < 2) throw Error(CYMG_ERR_INVALID, "a fiber needs at least 2 points");
        if (np > (uint32_t)kMaxFiberPoints)
            throw Error(CYMG_ERR_UNSUPPORTED, "fiber has more than CYMG_MAX_FIBER_POINTS points");
        if (!(segmentation > 0)) throw Error(CYMG_ERR_INVALID, "segmentation must be > 0");
        if (!(rigidity >= 0)) throw Error(CYMG_ERR_INVALID, "fiber:rigidity must be specified and >= 0");
        if (c.nsol) throw Error(CYMG_ERR_INVALID, "fibers are registered before Solids (simul_solve.cc:204-207)");
        if (c.h_blkoff.empty()) c.h_blkoff.assign(1, 0);
        if (mat_index) *mat_index = c.npts;
        c.h_np.push_back(np);
        c.h_idx.push_back(c.npts);
        c.h_seg.push_back(c.nseg);
        c.h_cut.push_back(segmentation);
        c.h_rig.push_back(rigidity);
        c.h_drag.insert(c.h_drag.end(), drag, drag + 3);
        c.h_ids.clear();
        const uint64_t bs = (uint64_t)c.dim * np;
        c.blk_total += bs * bs;
        c.h_blkoff.push_back(c.blk_total);
        c.npts += np;
        c.fib_pts = c.npts;
        c.nseg += np - 1;
        c.largest = std::max(c.largest, np);
        c.nfib += 1;
        c.layout_dirty = true;
        c.prepared = c.setup_done = c.pc_ready = c.solved = false;
}

int cymg_add_fiber(cymg_context* ctx, uint32_t np, double segmentation, double rigidity, uint32_t* mat_index)
{
    return guarded(ctx, [&](Context& c) {
        const double drag[3] = { 0.0, 0.0, 0.0 };
        add_fiber(c, np, segmentation, rigidity, drag, mat_index);
    });
}

int cymg_add_fiber_ex(cymg_context* ctx, const cymg_fiber_desc* d, uint32_t* mat_index)
{
    return guarded(ctx, [&](Context& c) {
        if (!d) throw Error(CYMG_ERR_INVALID, "null fiber description");
        if (!(d->hydro_radius0 >= 0) || !(d->hydro_radius1 >= 0))
            throw Error(CYMG_ERR_INVALID, "fiber:hydrodynamic_radius must be >= 0 (0: the context's value)");
        // Fiber::dragCoefficientSurface (fiber.cc:405-406)
        if (d->surface_effect && !(d->cylinder_height > 0))
            throw Error(CYMG_ERR_INVALID, "fiber:surface_effect[1] (height above surface) must set and > 0!");
        const double drag[3] = { d->hydro_radius0, d->hydro_radius1, d->surface_effect ? d->cylinder_height : 0.0 };
        add_fiber(c, d->nb_points, d->segmentation, d->rigidity, drag, mat_index);
    });
}

int cymg_set_identities(cymg_context* ctx, uint32_t n, const uint32_t* ids)
{
    return guarded(ctx, [&](Context& c) {
        if (n != c.nmec() || (n && !ids)) throw Error(CYMG_ERR_INVALID, "one identity per registered Mecable");
        std::vector<uint32_t> v(ids, ids + n);
        std::vector<uint32_t> sorted = v;
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            throw Error(CYMG_ERR_INVALID, "identities must be unique");
        c.h_ids = v;
        if (!c.layout_dirty) c.h_prev_ids = v;   // the uploaded registration is the one named
    });
}

int cymg_add_solid(cymg_context* ctx, uint32_t np, const double* radius, const double* points, int reshape_timer,
                   uint32_t* mat_index)
{
    return guarded(ctx, [&](Context& c) {
        if (np < 1 || !radius || !points) throw Error(CYMG_ERR_INVALID, "a Solid needs points and radii");
        if (np > (uint32_t)CYMG_MAX_SOLID_POINTS)
            throw Error(CYMG_ERR_UNSUPPORTED, "Solid has more than CYMG_MAX_SOLID_POINTS points");
        if (c.h_sys_begin.size() > 2) throw Error(CYMG_ERR_UNSUPPORTED, "Solids in a batch of independent systems");
        for (uint32_t p = 0; p < np; ++p)
            if (!(radius[p] >= 0)) throw Error(CYMG_ERR_INVALID, "solid:sphere's radius should be >= 0");
        const int D = c.dim;
        if (mat_index) *mat_index = c.npts;
        c.h_sol_np.push_back(np);
        c.h_sol_idx.push_back(c.npts);
        c.h_sol_rad.insert(c.h_sol_rad.end(), radius, radius + np);
        // Solid::fixShape (solid.cc:370-400) with PointSet::calculateMomentum (point_set.cc:288-317)
        double cc[3] = { 0, 0, 0 }, pp[3] = { 0, 0, 0 };
        for (int d = 0; d < D; ++d) {
            cc[d] = points[d];
            pp[d] = points[d] * points[d];
        }
        for (uint32_t p = 1; p < np; ++p)
            for (int d = 0; d < D; ++d) {
                cc[d] += points[D * p + d];
                pp[d] += points[D * p + d] * points[D * p + d];
            }
        const double N = (double)np;
        for (int d = 0; d < D; ++d) cc[d] *= 1.0 / N;
        double sq = 0.0;
        for (int d = 0; d < D; ++d) sq += pp[d] - N * cc[d] * cc[d];
        c.h_sol_shape_sqr.push_back(sq);
        for (uint32_t p = 0; p < np; ++p)
            for (int d = 0; d < D; ++d) c.h_sol_shape.push_back(points[D * p + d] - cc[d]);
        c.h_sol_timer.push_back(reshape_timer);
        c.npts += np;
        c.nsol += 1;
        c.h_ids.clear();
        c.layout_dirty = true;
        c.prepared = c.setup_done = c.pc_ready = c.solved = false;
    });
}

int cymg_get_solid_timers(cymg_context* ctx, uint32_t nsol, int* timers)
{
    return guarded(ctx, [&](Context& c) {
        if (nsol != c.nsol || (nsol && !timers)) throw Error(CYMG_ERR_INVALID, "nsol must match the Solids registered");
        if (!nsol) return;
        if (c.layout_dirty) {
            std::copy(c.h_sol_timer.begin(), c.h_sol_timer.end(), timers);
            return;
        }
        CYMG_HIP(hipMemcpyAsync(timers, c.d_sol_timer.get(), nsol * sizeof(int), hipMemcpyDeviceToHost, c.stream));
        CYMG_HIP(hipStreamSynchronize(c.stream));
    });
}

int cymg_set_systems(cymg_context* ctx, uint32_t nsys, const uint32_t* first_fiber)
{
    return guarded(ctx, [&](Context& c) {
        if (nsys < 1 || nsys > 4096) throw Error(CYMG_ERR_UNSUPPORTED, "between 1 and 4096 systems");
        if (nsys > 1 && c.nsol) throw Error(CYMG_ERR_UNSUPPORTED, "Solids in a batch of independent systems");
        std::vector<uint32_t> sb(first_fiber, first_fiber + nsys);
        sb.push_back(c.nfib);
        if (sb[0] != 0) throw Error(CYMG_ERR_INVALID, "the first system must start at fiber 0");
        for (uint32_t s = 0; s < nsys; ++s)
            if (sb[s + 1] <= sb[s]) throw Error(CYMG_ERR_INVALID, "systems must be non-empty and in order");
        c.h_sys_begin = sb;
        c.layout_dirty = true;
        c.prepared = c.setup_done = c.pc_ready = c.solved = false;
    });
}

int cymg_get_system_reports(cymg_context* ctx, uint32_t nsys, cymg_solve_report* out)
{
    return guarded(ctx, [&](Context& c) {
        if (!out || nsys != c.sys_reports.size()) throw Error(CYMG_ERR_INVALID, "nsys must match the last solve");
        for (uint32_t s = 0; s < nsys; ++s) out[s] = c.sys_reports[s];
    });
}

int cymg_nb_points(const cymg_context* ctx, uint32_t* nb)
{
    if (!ctx || !nb) return CYMG_ERR_INVALID;
    *nb = ctx->c.npts;
    return CYMG_OK;
}

int cymg_prepare(cymg_context* ctx, const double* points)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.nmec()) throw Error(CYMG_ERR_INVALID, "no Mecable registered");
        if (c.nsol && !c.layout_dirty && c.d_status.get()) {
            // Solid::rescale's ABORT_NOW (solid.cc:449-453) in the previous getPoints
            int st = 0;
            CYMG_HIP(hipMemcpyAsync(&st, c.d_status.get() + 2, sizeof(int), hipMemcpyDeviceToHost, c.stream));
            CYMG_HIP(hipStreamSynchronize(c.stream));
            if (st) throw Error(CYMG_ERR_INVALID, "Cannot rescale the Solid, because it has collapsed!");
        }
        const bool fresh = c.layout_dirty;
        if (!points && !(fresh ? c.prev_layout : c.have_solution))
            throw Error(CYMG_ERR_INVALID, "no points on the device: pass the Mecable points");
        if (c.layout_dirty) upload_layout(c, points == nullptr);
        if (points) {
            CYMG_HIP(hipMemcpyAsync(c.d_pos.get(), points, (size_t)c.n() * 8, hipMemcpyHostToDevice, c.stream));
            c.have_solution = true;
        } else if (fresh && !c.pos_complete) {
            c.have_solution = false;
            throw Error(CYMG_ERR_INVALID, "the device does not hold the points of every Mecable registered: "
                                          "pass the Mecable points (or give identities, cymg_set_identities)");
        }
        c.hB.clear(); c.hC.clear(); c.hBas.clear();
        c.marks.clear();
        c.h_def.clear();
        launch_prepare(c);
        c.prepared = true;
        c.setup_done = false;
        c.pc_ready = false;
        c.solved = false;
    });
}

// ---- assembly, in call order -------------------------------------------------
// Products and sums are rounded as written (no FMA contraction), the
// arithmetic of the reference source.
#pragma clang fp contract(off)

int cymg_add_pure_force(cymg_context* ctx, uint32_t a, const double* force)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_point(c, a);
        for (int d = 0; d < c.dim; ++d) addBas(c, c.dim * a + d, force[d]);
    });
}

int cymg_inter_link_ee(cymg_context* ctx, uint32_t a, uint32_t b, double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_point(c, a); check_point(c, b);
        if (!(w >= 0)) throw Error(CYMG_ERR_INVALID, "weight must be >= 0");
        same_system(c, { a, b });
        if (mecable_of(c, a) == mecable_of(c, b) && (a == b || a == b + 1 || a + 1 == b)) return;
        addB(c, a, a, -w);
        addB(c, a, b, +w);
        addB(c, b, b, -w);
    });
}

int cymg_inter_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b, double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_point(c, a1); check_point(c, a2); check_point(c, b);
        if (!(w >= 0)) throw Error(CYMG_ERR_INVALID, "weight must be >= 0");
        same_system(c, { a1, a2, b });
        if (mecable_of(c, a1) == mecable_of(c, b) && (a1 == b || a2 == b)) return;
        const double cc[3] = { 1.0 - ca, ca, -1.0 };
        const double cw[3] = { w * cc[0], w * cc[1], -w };
        const uint32_t inx[3] = { a1, a2, b };
        for (int k = 0; k < 3; ++k)
            for (int l = k; l < 3; ++l) addB(c, inx[k], inx[l], -(cc[k] * cw[l]));
    });
}

int cymg_inter_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                       double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_point(c, a1); check_point(c, a2); check_point(c, b1); check_point(c, b2);
        if (!(w >= 0)) throw Error(CYMG_ERR_INVALID, "weight must be >= 0");
        same_system(c, { a1, a2, b1, b2 });
        if (mecable_of(c, a1) == mecable_of(c, b1) && (a1 == b1 || a1 == b2 || a2 == b1 || a2 == b2)) return;
        const double cc[4] = { 1.0 - ca, ca, -(1.0 - cb), -cb };
        const double cw[4] = { w * cc[0], w * cc[1], w * cc[2], w * cc[3] };
        const uint32_t inx[4] = { a1, a2, b1, b2 };
        for (int j = 0; j < 4; ++j)
            for (int i = j; i < 4; ++i) addB(c, inx[i], inx[j], -(cc[j] * cw[i]));
    });
}

// A position-dependent emitter is evaluated on the device when the streams are
// built (emit.hip); consecutive calls share one mark, placed after the host
// entries recorded before the first of them.
static void defer(Context& c, DeferredCall q)
{
    uint32_t s[3];
    deferred_slots(q.kind, c.dim, s);
    const size_t pos[3] = { c.hB.key.size(), c.hC.key.size(), c.hBas.key.size() };
    if (c.marks.empty() || c.marks.back().kind != 4 || c.marks.back().posB != pos[0] ||
        c.marks.back().posC != pos[1] || c.marks.back().posBas != pos[2]) {
        EmitterMark mk{ 4, 0.0, 0.0, pos[0], pos[1], pos[2] };
        mk.first = (uint32_t)c.h_def.size();
        c.marks.push_back(mk);
    }
    EmitterMark& mk = c.marks.back();
    for (int w = 0; w < 3; ++w) {
        if (mk.slots[w] + s[w] > 0xFFFFFFFFull) throw Error(CYMG_ERR_UNSUPPORTED, "too many deferred interactions");
        q.off[w] = (uint32_t)mk.slots[w];
        mk.slots[w] += s[w];
    }
    c.h_def.push_back(q);
    ++mk.count;
}

static DeferredCall call_of(uint32_t kind, std::initializer_list<uint32_t> pts, double ca, double cb, double len,
                            double w, const double* g = nullptr, int ng = 0)
{
    DeferredCall q{};
    q.kind = kind;
    int k = 0;
    for (uint32_t p : pts) q.p[k++] = p;
    q.ca = ca;
    q.cb = cb;
    q.len = len;
    q.w = w;
    for (int d = 0; d < ng; ++d) q.g[d] = g[d];
    return q;
}

// PointInterpolated: two points of one Mecable (point_interpolated.h:33-60)
static void check_interpolated(const Context& c, uint32_t p1, uint32_t p2)
{
    check_point(c, p1);
    check_point(c, p2);
    if (p1 == p2 || mecable_of(c, p1) != mecable_of(c, p2))
        throw Error(CYMG_ERR_INVALID, "an interpolated point needs two distinct points of one Mecable");
}

static void check_weight(double w)
{
    if (!(w >= 0)) throw Error(CYMG_ERR_INVALID, "weight must be >= 0");
}

// PointInterpolated::overlapping (point_interpolated.cc:21-33)
static bool overlapping(const Context& c, uint32_t a1, uint32_t a2, uint32_t b)
{
    return mecable_of(c, a1) == mecable_of(c, b) && (a1 == b || a2 == b);
}

static bool overlapping(const Context& c, uint32_t a1, uint32_t a2, uint32_t b1, uint32_t b2)
{
    return mecable_of(c, a1) == mecable_of(c, b1) && (a1 == b1 || a1 == b2 || a2 == b1 || a2 == b2);
}

int cymg_inter_long_link_ee(cymg_context* ctx, uint32_t a, uint32_t b, double len, double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_point(c, a); check_point(c, b);
        check_weight(w);
        same_system(c, { a, b });
        if (mecable_of(c, a) == mecable_of(c, b) && (a == b || a == b + 1 || a + 1 == b)) return;
        defer(c, call_of(kDefLongLinkEE, { a, b }, 0.0, 0.0, len, w));
    });
}

int cymg_inter_clamp_e(cymg_context* ctx, uint32_t a, const double* g, double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_point(c, a);
        addB(c, a, a, -w);
        for (int d = 0; d < c.dim; ++d) addBas(c, c.dim * a + d, w * g[d]);
    });
}

int cymg_inter_clamp_i(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, const double* g, double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_point(c, a1); check_point(c, a2);
        const double c2 = ca, c2w = w * c2;
        const double c1 = 1.0 - ca, c1w = w * c1;
        addB(c, a1, a1, -(c1w * c1));
        addB(c, a1, a2, -(c2w * c1));
        addB(c, a2, a2, -(c2w * c2));
        for (int d = 0; d < c.dim; ++d) {
            addBas(c, c.dim * a1 + d, c1w * g[d]);
            addBas(c, c.dim * a2 + d, c2w * g[d]);
        }
    });
}

int cymg_inter_long_clamp(cymg_context* ctx, uint32_t a, const double* center, double len, double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_point(c, a);
        if (!center) throw Error(CYMG_ERR_INVALID, "null center");
        defer(c, call_of(kDefLongClamp, { a }, 0.0, 0.0, len, w, center, c.dim));
    });
}

int cymg_inter_plane_e(cymg_context* ctx, uint32_t a, const double* dir, const double* g, double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_point(c, a);
        const int D = c.dim;
        const uint32_t inx = D * a;
        double gd = 0;
        for (int d = 0; d < D; ++d) gd += g[d] * dir[d];
        const double pr = gd * w;
        for (int i = 0; i < D; ++i) {
            for (int j = i; j < D; ++j) addC(c, inx + i, inx + j, -(dir[i] * dir[j] * w));
            addBas(c, inx + i, pr * dir[i]);
        }
    });
}

int cymg_add_pure_force_i(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, const double* force)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_interpolated(c, a1, a2);
        if (!force) throw Error(CYMG_ERR_INVALID, "null force");
        const double a = ca, b = 1.0 - ca;
        for (int d = 0; d < c.dim; ++d) {
            addBas(c, c.dim * a1 + d, b * force[d]);
            addBas(c, c.dim * a2 + d, a * force[d]);
        }
    });
}

int cymg_inter_plane_i(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, const double* dir, const double* g,
                       double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_interpolated(c, a1, a2);
        if (!dir || !g) throw Error(CYMG_ERR_INVALID, "null vector");
        const int D = c.dim;
        const double aa = ca, bb = 1.0 - ca;
        const double bbbb = bb * bb, aaaa = aa * aa, aabb = aa * bb;
        const uint32_t i1 = D * a1, i2 = D * a2;
        for (int x = 0; x < D; ++x) {
            for (int y = x; y < D; ++y) {
                const double P = w * dir[x] * dir[y];
                addC(c, i1 + x, i1 + y, -(bbbb * P));
                addC(c, i2 + x, i2 + y, -(aaaa * P));
                addC(c, i1 + x, i2 + y, -(aabb * P));
                if (x != y) addC(c, i1 + y, i2 + x, -(aabb * P));
            }
        }
        double gd = g[0] * dir[0];
        for (int d = 1; d < D; ++d) gd = gd + g[d] * dir[d];
        const double pr = w * gd;
        for (int x = 0; x < D; ++x) {
            addBas(c, i1 + x, pr * bb * dir[x]);
            addBas(c, i2 + x, pr * aa * dir[x]);
        }
    });
}

// interpolated-to-exact and interpolated-to-interpolated emitters evaluated
// from the device points
static int defer_ie(cymg_context* ctx, uint32_t kind, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len,
                    double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_interpolated(c, a1, a2); check_point(c, b);
        check_weight(w);
        same_system(c, { a1, a2, b });
        if (overlapping(c, a1, a2, b)) return;
        defer(c, call_of(kind, { a1, a2, b }, ca, 0.0, len, w));
    });
}

static int defer_ii(cymg_context* ctx, uint32_t kind, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                    double cb, double len, double w, const double* g = nullptr, bool plane_only = false)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_interpolated(c, a1, a2); check_interpolated(c, b1, b2);
        check_weight(w);
        if (plane_only && c.dim != 2) {
            throw Error(CYMG_ERR_UNSUPPORTED, kind == kDefTorque ? "interTorque2D exists in 2D only"
                                                                  : "interSideSideLink() was not implemented in 3D");
        }
        same_system(c, { a1, a2, b1, b2 });
        if (overlapping(c, a1, a2, b1, b2)) return;
        defer(c, call_of(kind, { a1, a2, b1, b2 }, ca, cb, len, w, g, g ? 2 : 0));
    });
}

int cymg_inter_long_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len, double w)
{
    return defer_ie(ctx, kDefLongLinkIE, a1, a2, ca, b, len, w);
}

int cymg_inter_long_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                            double cb, double len, double w)
{
    return defer_ii(ctx, kDefLongLinkII, a1, a2, ca, b1, b2, cb, len, w);
}

int cymg_inter_side_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len, double w)
{
    return defer_ie(ctx, kDefSideLinkIE, a1, a2, ca, b, len, w);
}

int cymg_inter_side_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                            double cb, double len, double w)
{
    return defer_ii(ctx, kDefSideLinkII, a1, a2, ca, b1, b2, cb, len, w);
}

int cymg_inter_side_side_link(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                              double cb, double len, double w)
{
    return defer_ii(ctx, kDefSideSideLink, a1, a2, ca, b1, b2, cb, len, w, nullptr, true);
}

int cymg_inter_sliding_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b, double w)
{
    return defer_ie(ctx, kDefSlidingLinkIE, a1, a2, ca, b, 0.0, w);
}

int cymg_inter_sliding_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                               double cb, double w)
{
    return defer_ii(ctx, kDefSlidingLinkII, a1, a2, ca, b1, b2, cb, 0.0, w);
}

int cymg_inter_side_sliding_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len,
                                    double w)
{
    return defer_ie(ctx, kDefSideSlidingIE, a1, a2, ca, b, len, w);
}

int cymg_inter_side_sliding_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1,
                                    uint32_t b2, double cb, double len, double w)
{
    return defer_ii(ctx, kDefSideSlidingII, a1, a2, ca, b1, b2, cb, len, w);
}

int cymg_inter_torque(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                      double cosinus, double sinus, double w)
{
    const double cs[2] = { cosinus, sinus };
    return defer_ii(ctx, kDefTorque, a1, a2, ca, b1, b2, cb, 0.0, w, cs, true);
}

int cymg_inter_side_clamp(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, const double* g, double len,
                          double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_interpolated(c, a1, a2);
        check_weight(w);
        if (!g) throw Error(CYMG_ERR_INVALID, "null position");
        defer(c, call_of(kDefSideClamp, { a1, a2 }, ca, 0.0, len, w, g, c.dim));
    });
}

int cymg_inter_sliding_clamp(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, const double* g, double w)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c); check_interpolated(c, a1, a2);
        check_weight(w);
        if (!g) throw Error(CYMG_ERR_INVALID, "null position");
        defer(c, call_of(kDefSlidingClamp, { a1, a2 }, ca, 0.0, 0.0, w, g, c.dim));
    });
}

int cymg_set_steric(cymg_context* ctx, const cymg_steric_params* params, uint32_t nmec, const int* steric,
                    const double* radius, const double* range)
{
    return guarded(ctx, [&](Context& c) {
        if (!params || (nmec && (!steric || !radius || !range))) throw Error(CYMG_ERR_INVALID, "null argument");
        if (!(params->stiff_push >= 0) || !(params->stiff_pull >= 0))
            throw Error(CYMG_ERR_INVALID, "steric stiffness must be >= 0");
        c.st_prm = *params;
        c.h_st_flag.assign(steric, steric + nmec);
        c.h_st_rad.assign(radius, radius + nmec);
        c.h_st_rng.assign(range, range + nmec);
        c.st_set = true;
        c.st_gen = ~0ull;
    });
}

int cymg_steric_interactions(cymg_context* ctx)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c);
        if (!c.st_set) throw Error(CYMG_ERR_INVALID, "call cymg_set_steric first");
        if (c.nsys > 1) throw Error(CYMG_ERR_UNSUPPORTED, "steric interactions in a batch of independent systems");
        steric_interactions(c);
    });
}

int cymg_set_frame_records(cymg_context* ctx, uint32_t nmec, const cymg_frame_record* records)
{
    return guarded(ctx, [&](Context& c) {
        if (nmec && !records) throw Error(CYMG_ERR_INVALID, "null argument");
        c.h_frame_rec.assign(records, records + nmec);
        ++c.frame_rec_gen;
    });
}

int cymg_write_frame(cymg_context* ctx, const cymg_frame_params* params, void* out, size_t capacity, size_t* size)
{
    return guarded(ctx, [&](Context& c) {
        if (!params || !size) throw Error(CYMG_ERR_INVALID, "null argument");
        need_prepared(c);
        write_frame(c, *params, out, capacity, size);
    });
}

int cymg_confine_sphere(cymg_context* ctx, double radius, double stiffness)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c);
        if (!(radius > 0) || !(stiffness >= 0)) throw Error(CYMG_ERR_INVALID, "invalid confinement");
        c.marks.push_back({ 1, radius, stiffness, c.hB.key.size(), c.hC.key.size(), c.hBas.key.size() });
    });
}

int cymg_set_couples(cymg_context* ctx, uint32_t n, const uint32_t* fa, const double* aa, const uint32_t* fb,
                     const double* ab, const double* w)
{
    return guarded(ctx, [&](Context& c) {
        for (uint32_t k = 0; k < n; ++k) {
            if (fa[k] >= c.nfib || fb[k] >= c.nfib) throw Error(CYMG_ERR_INVALID, "couple on an unknown fiber");
            if (!(w[k] >= 0)) throw Error(CYMG_ERR_INVALID, "couple stiffness must be >= 0");
            if (system_of_fiber(c, fa[k]) != system_of_fiber(c, fb[k]))
                throw Error(CYMG_ERR_INVALID, "couple links two independent systems");
        }
        c.ncouples = n;
        ++c.couples_gen;
        const size_t m = n ? n : 1;
        c.cpl_fa.resize(m); c.cpl_fb.resize(m); c.cpl_aa.resize(m); c.cpl_ab.resize(m); c.cpl_w.resize(m);
        c.cpl_state.resize(m);
        // both hands attached, passive (a crosslinker)
        CYMG_HIP(hipMemsetAsync(c.cpl_state.get(), 3, m, c.stream));
        if (n) {
            CYMG_HIP(hipMemcpyAsync(c.cpl_fa.get(), fa, n * 4, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.cpl_fb.get(), fb, n * 4, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.cpl_aa.get(), aa, n * 8, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.cpl_ab.get(), ab, n * 8, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.cpl_w.get(), w, n * 8, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

static void d2h(Context& c, void* dst, const void* src, size_t bytes);

int cymg_set_couple_state(cymg_context* ctx, const uint8_t* state)
{
    return guarded(ctx, [&](Context& c) {
        if (!state) throw Error(CYMG_ERR_INVALID, "null array");
        for (uint32_t k = 0; k < c.ncouples; ++k)
            if (state[k] > 15) throw Error(CYMG_ERR_INVALID, "couple state has bits other than 1, 2, 4, 8");
        if (c.ncouples) {
            CYMG_HIP(hipMemcpyAsync(c.cpl_state.get(), state, c.ncouples, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipStreamSynchronize(c.stream));
        }
        ++c.couples_gen;
    });
}

int cymg_nearby_segments(cymg_context* ctx, uint32_t nq, const double* places, const uint32_t* exclude,
                         double range, int inclusive, uint32_t* counts, uint32_t capacity, uint32_t* fiber,
                         uint32_t* point, double* abscissa, double* dist_sqr, uint32_t* total)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.prepared) throw Error(CYMG_ERR_INVALID, "no points: cymg_prepare first");
        if (!(range >= 0)) throw Error(CYMG_ERR_INVALID, "range must be >= 0");
        if (!total || (nq && (!places || !counts))) throw Error(CYMG_ERR_INVALID, "null argument");
        if (capacity && (!fiber || !point || !abscissa || !dist_sqr)) throw Error(CYMG_ERR_INVALID, "null output");
        nearby_segments(c, nq, places, exclude, range, inclusive, counts, capacity, fiber, point, abscissa,
                        dist_sqr, total);
    });
}

int cymg_nearby_segments_noalloc(cymg_context* ctx, uint32_t nq, const double* places, const uint32_t* exclude,
                                 double range, int inclusive, uint32_t* counts, uint32_t capacity, uint32_t* fiber,
                                 uint32_t* point, double* abscissa, double* dist_sqr, uint32_t* total)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.prepared) throw Error(CYMG_ERR_INVALID, "no points: cymg_prepare first");
        if (!(range >= 0)) throw Error(CYMG_ERR_INVALID, "range must be >= 0");
        if (!total || (nq && (!places || !counts))) throw Error(CYMG_ERR_INVALID, "null argument");
        if (capacity && (!fiber || !point || !abscissa || !dist_sqr)) throw Error(CYMG_ERR_INVALID, "null output");
        nearby_segments_noalloc(c, nq, places, exclude, range, inclusive, counts, capacity, fiber, point, abscissa,
                                dist_sqr, total);
    });
}

int cymg_get_couples(cymg_context* ctx, double* aa, double* ab, uint8_t* state)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.ncouples) return;
        if (aa) d2h(c, aa, c.cpl_aa.get(), (size_t)c.ncouples * 8);
        if (ab) d2h(c, ab, c.cpl_ab.get(), (size_t)c.ncouples * 8);
        if (state) d2h(c, state, c.cpl_state.get(), c.ncouples);
    });
}

int cymg_step_couples(cymg_context* ctx, const cymg_motor_props* mp)
{
    return guarded(ctx, [&](Context& c) {
        if (!mp) throw Error(CYMG_ERR_INVALID, "null motor properties");
        if (!(mp->stall_force > 0)) throw Error(CYMG_ERR_INVALID, "motor:stall_force must be > 0");
        if (!c.prepared) throw Error(CYMG_ERR_INVALID, "no points: cymg_prepare first");
        // MotorProp::complete (motor_prop.cc:44-66)
        MotorStep ms;
        ms.max_speed_dt = c.prm.time_step * mp->max_speed;
        const double abs_speed_dt = fabs(ms.max_speed_dt);
        ms.var_speed_dt = abs_speed_dt / mp->stall_force;
        if (mp->max_speed > 0) {
            ms.min_dabs = 0;
            ms.max_dabs = 2 * ms.max_speed_dt;
        } else {
            ms.min_dabs = 2 * ms.max_speed_dt;
            ms.max_dabs = 0;
        }
        ms.limit_speed = mp->limit_speed != 0;
        ms.hold_end = mp->hold_growing_end != 0;
        launch_step_couples(c, ms);
        ++c.couples_gen;   // the hands' segments, hence the couple keys, may change
    });
}

int cymg_couples_interactions(cymg_context* ctx)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c);
        c.marks.push_back({ 2, 0.0, 0.0, c.hB.key.size(), c.hC.key.size(), c.hBas.key.size() });
    });
}

int cymg_set_aster_links(cymg_context* ctx, uint32_t n, const int* kind, const uint32_t* s1, const uint32_t* s2,
                         const uint32_t* fiber, const double* ab, const double* w)
{
    return guarded(ctx, [&](Context& c) {
        if (n && (!kind || !s1 || !s2 || !fiber || !ab || !w)) throw Error(CYMG_ERR_INVALID, "null array");
        for (uint32_t k = 0; k < n; ++k) {
            if (kind[k] < 0 || kind[k] > 2) throw Error(CYMG_ERR_INVALID, "aster link kind must be 0, 1 or 2");
            if (fiber[k] >= c.nfib) throw Error(CYMG_ERR_INVALID, "aster link on an unknown fiber");
            if (s1[k] < c.fib_pts || s1[k] >= c.npts || (kind[k] == 2 && (s2[k] < c.fib_pts || s2[k] >= c.npts)))
                throw Error(CYMG_ERR_INVALID, "aster link on a point that is not a Solid point");
            if (!(w[k] >= 0)) throw Error(CYMG_ERR_INVALID, "aster stiffness must be >= 0");
        }
        c.nasters = n;
        ++c.asters_gen;
        const size_t m = n ? n : 1;
        c.ast_kind.resize(m); c.ast_s1.resize(m); c.ast_s2.resize(m); c.ast_f.resize(m);
        c.ast_ab.resize(m); c.ast_w.resize(m);
        if (n) {
            CYMG_HIP(hipMemcpyAsync(c.ast_kind.get(), kind, n * 4, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.ast_s1.get(), s1, n * 4, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.ast_s2.get(), s2, n * 4, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.ast_f.get(), fiber, n * 4, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.ast_ab.get(), ab, n * 8, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.ast_w.get(), w, n * 8, hipMemcpyHostToDevice, c.stream));
            CYMG_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int cymg_aster_interactions(cymg_context* ctx)
{
    return guarded(ctx, [&](Context& c) {
        need_prepared(c);
        c.marks.push_back({ 3, 0.0, 0.0, c.hB.key.size(), c.hC.key.size(), c.hBas.key.size() });
    });
}

// ---- solve ---------------------------------------------------------------------

int cymg_setup(cymg_context* ctx, const double* gauss, int on_device)
{
    return guarded(ctx, [&](Context& c) { do_setup(c, gauss, on_device); });
}

int cymg_iterate(cymg_context* ctx, const cymg_solve_params* p, cymg_solve_report* rep)
{
    return guarded(ctx, [&](Context& c) {
        if (!p) throw Error(CYMG_ERR_INVALID, "null parameters");
        do_iterate(c, p, rep);
        collect_timings(c);
    });
}

int cymg_apply(cymg_context* ctx)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.setup_done || !c.solved)
            throw Error(CYMG_ERR_INVALID, "nothing solved: cymg_iterate must succeed after the last setup");
        do_apply(c);
    });
}

int cymg_solve(cymg_context* ctx, const cymg_solve_params* p, const double* gauss, cymg_solve_report* rep)
{
    return guarded(ctx, [&](Context& c) {
        if (!p) throw Error(CYMG_ERR_INVALID, "null parameters");
        do_setup(c, gauss, p->gauss_on_device);
        do_iterate(c, p, rep);
        do_apply(c);
        collect_timings(c);
    });
}

// ---- operators -------------------------------------------------------------------

int cymg_multiply(cymg_context* ctx, const double* X, double* Y)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede cymg_multiply");
        const size_t n = c.n();
        c.k_v.resize(n); c.k_t.resize(n);
        CYMG_HIP(hipMemcpyAsync(c.k_v.get(), X, n * 8, hipMemcpyHostToDevice, c.stream));
        launch_matvec(c, c.k_v.get(), c.k_t.get());
        CYMG_HIP(hipMemcpyAsync(Y, c.k_t.get(), n * 8, hipMemcpyDeviceToHost, c.stream));
        CYMG_HIP(hipStreamSynchronize(c.stream));
    });
}

int cymg_multiply_device(cymg_context* ctx, const double* dX, double* dY)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede cymg_multiply");
        launch_matvec(c, dX, dY);
    });
}

int cymg_compute_preconditioner(cymg_context* ctx, int precision)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede the preconditioner");
        if (precision != CYMG_PRECOND_FP64 && precision != CYMG_PRECOND_FP32)
            throw Error(CYMG_ERR_INVALID, "precision must be 1 (FP64) or 2 (FP32)");
        launch_precond_build(c, precision);
        CYMG_HIP(hipStreamSynchronize(c.stream));
    });
}

int cymg_precondition(cymg_context* ctx, const double* X, double* Y)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.pc_ready) throw Error(CYMG_ERR_INVALID, "no preconditioner computed");
        const size_t n = c.n();
        c.k_v.resize(n); c.k_t.resize(n);
        CYMG_HIP(hipMemcpyAsync(c.k_v.get(), X, n * 8, hipMemcpyHostToDevice, c.stream));
        launch_precond_apply(c, c.k_v.get(), c.k_t.get());
        CYMG_HIP(hipMemcpyAsync(Y, c.k_t.get(), n * 8, hipMemcpyDeviceToHost, c.stream));
        CYMG_HIP(hipStreamSynchronize(c.stream));
    });
}

int cymg_precondition_device(cymg_context* ctx, const double* dX, double* dY)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.pc_ready) throw Error(CYMG_ERR_INVALID, "no preconditioner computed");
        launch_precond_apply(c, dX, dY);
    });
}

int cymg_compute_forces(cymg_context* ctx)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.prepared) throw Error(CYMG_ERR_INVALID, "cymg_prepare must precede cymg_compute_forces");
        build_streams(c);
        launch_forces(c, c.d_pts.get(), c.d_for.get(), false);
        launch_tensions(c);
        CYMG_HIP(hipStreamSynchronize(c.stream));
    });
}

// ---- export ------------------------------------------------------------------------

static void d2h(Context& c, void* dst, const void* src, size_t bytes)
{
    CYMG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    CYMG_HIP(hipStreamSynchronize(c.stream));
}

int cymg_get_vector(cymg_context* ctx, int which, double* out)
{
    return guarded(ctx, [&](Context& c) {
        const double* src = nullptr;
        switch (which) {
            case CYMG_VEC_PTS: src = c.d_pts.get(); break;
            case CYMG_VEC_SOL: src = c.d_sol.get(); break;
            case CYMG_VEC_BAS: src = c.d_bas.get(); break;
            case CYMG_VEC_RHS: src = c.d_rhs.get(); break;
            case CYMG_VEC_FOR: src = c.d_for.get(); break;
            case CYMG_VEC_PREV: src = c.d_prev_sol.get(); break;
            default: throw Error(CYMG_ERR_INVALID, "unknown vector");
        }
        if (!src) throw Error(CYMG_ERR_INVALID, "vector not allocated");
        d2h(c, out, src, (size_t)c.n() * 8);
    });
}

// Meca::dump (meca.cc:957-990) with dumpDiagonal (934-947) and dumpMatrix
// (883-903): the same files, formats and order, written into `directory`
// (NULL or "": the working directory, as the reference).  The matrix columns
// are the operator applied to the unit vectors on the device.
int cymg_dump(cymg_context* ctx, const char* directory)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede cymg_dump");
        const std::string dir = (directory && *directory) ? std::string(directory) + "/" : std::string();
        auto open = [&](const char* name, const char* mode) {
            FILE* f = std::fopen((dir + name).c_str(), mode);
            if (!f) throw Error(CYMG_ERR_INVALID, "cannot open " + dir + name);
            return f;
        };
        const size_t n = c.n();
        const int D = c.dim;
        FILE* f = open("ddim.txt", "w");
        std::fprintf(f, "%i\n", (int)n);
        std::fclose(f);
        // drag per point of every Mecable at its current points
        std::vector<double> pos(n), mob(c.nfib), geo((size_t)kSolidGeo * c.nsol);
        d2h(c, pos.data(), c.d_pos.get(), n * 8);
        if (c.nfib) d2h(c, mob.data(), c.d_mob.get(), c.nfib * 8);
        if (c.nsol) d2h(c, geo.data(), c.d_sol_geo.get(), geo.size() * 8);
        f = open("diagonal.txt", "w");
        for (uint32_t i = 0; i < c.nfib + c.nsol; ++i) {
            const bool fib = i < c.nfib;
            const uint32_t np = fib ? c.h_np[i] : c.h_sol_np[i - c.nfib];
            const uint32_t first = fib ? c.h_idx[i] : c.h_sol_idx[i - c.nfib];
            // RigidFiber::dragCoefficient = nbPoints / rfMobility; Solid: soDrag
            const double drag = (fib ? np / mob[i] : geo[(size_t)kSolidGeo * (i - c.nfib) + SG_DRAG]) / np;
            for (uint32_t p = 0; p < np; ++p)
                for (int d = 0; d < D; ++d)
                    std::fprintf(f, "%5u %16.8e %16.8e\n", i, pos[(size_t)D * (first + p) + d], drag);
        }
        std::fclose(f);
        std::vector<double> v(n);
        const std::pair<const char*, const double*> vecs[3] = {
            { "drhs.bin", c.d_rhs.get() }, { "dsol.bin", c.d_sol.get() }, { "dpts.bin", c.d_pts.get() } };
        for (const auto& kv : vecs) {
            if (kv.second) d2h(c, v.data(), kv.second, n * 8);
            else std::fill(v.begin(), v.end(), 0.0);
            f = open(kv.first, "wb");
            std::fwrite(v.data(), sizeof(double), n, f);
            std::fclose(f);
        }
        c.k_v.resize(n);
        c.k_t.resize(n);
        f = open("dmat.bin", "wb");
        const double one = 1.0;
        for (size_t i = 0; i < n; ++i) {
            CYMG_HIP(hipMemsetAsync(c.k_v.get(), 0, n * 8, c.stream));
            CYMG_HIP(hipMemcpyAsync(c.k_v.get() + i, &one, 8, hipMemcpyHostToDevice, c.stream));
            launch_matvec(c, c.k_v.get(), c.k_t.get());
            d2h(c, v.data(), c.k_t.get(), n * 8);
            std::fwrite(v.data(), sizeof(double), n, f);
        }
        std::fclose(f);
    });
}

int cymg_get_points(cymg_context* ctx, double* out)
{
    return guarded(ctx, [&](Context& c) { d2h(c, out, c.d_pos.get(), (size_t)c.n() * 8); });
}

int cymg_get_tensions(cymg_context* ctx, double* out)
{
    return guarded(ctx, [&](Context& c) { d2h(c, out, c.d_lag.get(), (size_t)c.nseg * 8); });
}

// dense block of Solid k from its low-rank form: I + G A, or I - G W
static void solid_dense_block(Context& c, uint32_t k, int inverse, double* out)
{
    const int D = c.dim, NG = D == 3 ? 6 : 3;
    const uint32_t np = c.h_sol_np[k], idx = c.h_sol_idx[k];
    const uint64_t bs = (uint64_t)D * np;
    std::vector<double> x(bs), F(6 * bs);
    d2h(c, x.data(), c.d_pts.get() + (size_t)D * idx, bs * 8);
    d2h(c, F.data(), (inverse ? c.d_sol_W.get() : c.d_sol_A.get()) + 6 * (size_t)D * (idx - c.fib_pts), 6 * bs * 8);
    const double sign = inverse ? -1.0 : 1.0;
    for (uint64_t j = 0; j < bs; ++j)
        for (uint64_t i = 0; i < bs; ++i) {
            const uint64_t p = i / D;
            const int d = (int)(i - D * p);
            const double* q = x.data() + D * p;
            double g[6] = { 0, 0, 0, 0, 0, 0 };
            g[d] = 1.0;
            if (D == 3) {
                g[3] = d == 0 ? 0.0 : (d == 1 ? -q[2] : q[1]);
                g[4] = d == 0 ? q[2] : (d == 1 ? 0.0 : -q[0]);
                g[5] = d == 0 ? -q[1] : (d == 1 ? q[0] : 0.0);
            } else {
                g[2] = d == 0 ? -q[1] : q[0];
            }
            double v = 0.0;
            for (int e = 0; e < NG; ++e) v += g[e] * F[6 * j + e];
            out[i + bs * j] = (i == j ? 1.0 : 0.0) + sign * v;
        }
}

int cymg_get_block(cymg_context* ctx, uint32_t m, int inverse, double* out)
{
    return guarded(ctx, [&](Context& c) {
        if (m >= c.nmec()) throw Error(CYMG_ERR_INVALID, "no such Mecable");
        if (m >= c.nfib) {
            const uint32_t k = m - c.nfib;
            if (inverse) {
                if (!c.pc_ready) throw Error(CYMG_ERR_INVALID, "no preconditioner computed");
                int use = 0;
                d2h(c, &use, c.d_sol_useW.get() + k, sizeof(int));
                if (!use) throw Error(CYMG_ERR_INVALID, "block not factorised (identity used)");
            } else {
                if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede cymg_get_block");
                launch_solid_pc_build(c);
            }
            solid_dense_block(c, k, inverse, out);
            return;
        }
        const uint64_t bs = (uint64_t)c.dim * c.h_np[m];
        const uint64_t off = c.h_blkoff[m];
        if (inverse) {
            if (!c.pc_ready) throw Error(CYMG_ERR_INVALID, "no preconditioner computed");
            int use = 0;
            d2h(c, &use, c.d_useblk.get() + m, sizeof(int));
            if (!use) throw Error(CYMG_ERR_INVALID, "block not factorised (identity used)");
            if (c.pc_precision == CYMG_PRECOND_FP32) {
                std::vector<float> tmp(bs * bs);
                d2h(c, tmp.data(), c.d_binv32.get() + off, bs * bs * 4);
                for (uint64_t i = 0; i < bs * bs; ++i) out[i] = tmp[i];
            } else {
                d2h(c, out, c.d_binv.get() + off, bs * bs * 8);
            }
        } else {
            if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede cymg_get_block");
            DevBuf<double> tmp;
            tmp.resize(c.blk_total);
            launch_precond_build(c, CYMG_PRECOND_FP64, tmp.get());
            d2h(c, out, tmp.get() + off, bs * bs * 8);
        }
    });
}

int cymg_get_matrix(cymg_context* ctx, int which, uint32_t* nnz, uint32_t* rows, uint32_t* cols, double* vals)
{
    return guarded(ctx, [&](Context& c) {
        if (!nnz) throw Error(CYMG_ERR_INVALID, "null nnz");
        DevSym& m = which ? c.mC : c.mB;
        *nnz = 0;
        if (!m.used) return;
        uint32_t nr = 0;
        d2h(c, &nr, m.nruns.get(), 4);
        std::vector<uint64_t> k(nr);
        std::vector<double> v(nr);
        d2h(c, k.data(), m.low_key.get(), nr * 8);
        d2h(c, v.data(), m.low_val.get(), nr * 8);
        // the reference creates the diagonal element of every column it touches
        // (matsparsesym1.cc:154-174); the device CSR leaves such zeros implicit,
        // the export restores them
        uint32_t cnt = 0;
        uint32_t prev_col = kInvalid32;
        for (uint32_t i = 0; i < nr; ++i) {
            if (k[i] == kInvalidKey) continue;
            const uint32_t col = (uint32_t)(k[i] >> 32), row = (uint32_t)k[i];
            if (col != prev_col && row != col) {
                if (rows) { rows[cnt] = col; cols[cnt] = col; vals[cnt] = 0.0; }
                ++cnt;
            }
            prev_col = col;
            if (rows) { rows[cnt] = row; cols[cnt] = col; vals[cnt] = v[i]; }
            ++cnt;
        }
        *nnz = cnt;
    });
}

int cymg_set_option(cymg_context* ctx, int option, int value)
{
    return guarded(ctx, [&](Context& c) {
        if (option == CYMG_OPT_KRYLOV_LAUNCHES && (value == 2 || value == 4)) c.krylov_fused = value == 2;
        else if (option == CYMG_OPT_BLOCK_INVERSE && value >= 0 && value <= 3) c.pc_method = value;
        else if (option == CYMG_OPT_GRAPHS && (value == 0 || value == 1)) c.use_graphs = value;
        else if (option == CYMG_OPT_TILE_POINTS && value >= 0 && value <= kTileMaxPoints) {
            c.tile_points = (uint32_t)value;
            if (!c.layout_dirty && c.nfib) build_tiles(c);
        }
        else if (option == CYMG_OPT_GUESS_GUARD && (value == 0 || value == 1)) c.mlp.guard = value;
        else if (option == CYMG_OPT_FP32_APPLY && (value == 0 || value == 1)) c.fp32_apply = value;
        else if (option == CYMG_OPT_BLOCK_STAGE && value >= 0 && value <= 2) c.block_stage = value;
        else if (option == CYMG_OPT_TILE_THREADS && (value == 256 || value == 512)) {
            c.tile_threads = (uint32_t)value;
            if (!c.layout_dirty && c.nfib) build_tiles(c);
        }
        else throw Error(CYMG_ERR_INVALID, "unknown option or value");
        c.last_iters[0] = c.last_iters[1] = c.last_iters[2] = 0;
    });
}

int cymg_get_option(cymg_context* ctx, int option, int* value)
{
    return guarded(ctx, [&](Context& c) {
        if (!value) throw Error(CYMG_ERR_INVALID, "null value");
        if (option == CYMG_OPT_KRYLOV_LAUNCHES) *value = c.krylov_fused ? 2 : 4;
        else if (option == CYMG_OPT_BLOCK_INVERSE) *value = c.pc_method;
        else if (option == CYMG_OPT_GRAPHS) *value = c.use_graphs;
        else if (option == CYMG_OPT_TILE_POINTS) *value = (int)c.tile_points;
        else if (option == CYMG_OPT_TILE_THREADS) *value = tile_threads_of(c);
        else if (option == CYMG_OPT_GUESS_GUARD) *value = c.mlp.guard;
        else if (option == CYMG_OPT_FP32_APPLY) *value = c.fp32_apply;
        else if (option == CYMG_OPT_BLOCK_STAGE) *value = c.block_stage;
        else throw Error(CYMG_ERR_INVALID, "unknown option");
    });
}

int cymg_autotune(cymg_context* ctx, int reps, int* tile_points, int* block_inverse, double* launch_ms)
{
    return guarded(ctx, [&](Context& c) {
        if (reps <= 0) throw Error(CYMG_ERR_INVALID, "reps must be > 0");
        if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede cymg_autotune");
        c.solved = false;   // the timed launches run on the solver's vectors
        const int precision = c.pc_ready ? c.pc_precision : CYMG_PRECOND_FP64;
        // block inverse: the fastest of Gauss-Jordan in registers, in LDS and
        // blocked on the matrix cores (all produce the inverse the iteration reads)
        const int methods[3] = { 0, 2, 3 };
        double best_build = 1e300;
        int best_method = c.pc_method;
        for (int m : methods) {
            c.pc_method = m;
            launch_precond_build(c, precision);
            CYMG_HIP(hipEventRecord(c.ev[8], c.stream));
            for (int i = 0; i < 3; ++i) launch_precond_build(c, precision);
            CYMG_HIP(hipEventRecord(c.ev[9], c.stream));
            CYMG_HIP(hipEventSynchronize(c.ev[9]));
            float ms = 0;
            CYMG_HIP(hipEventElapsedTime(&ms, c.ev[8], c.ev[9]));
            if (ms < best_build) { best_build = ms; best_method = m; }
        }
        c.pc_method = best_method;
        launch_precond_build(c, precision);
        // tile shape of the two-launch iteration: the target points per tile
        // with the shortest launch on the current system
        int best_tile = (int)c.tile_points, best_threads = (int)c.tile_threads;
        double best = 1e300;
        if ((!c.nsol || c.solid_tiles) && tiles_fit(c)) {
            const int cands[] = { 0, 16, 24, 32, 48, 64, 85, 96, 128, 170 };
            for (int nt : { 512, 256 }) {
                if (nt == 256 && c.largest > (uint32_t)tile_max_points(c, 256)) continue;
                for (int t : cands) {
                    if (t > tile_max_points(c, nt)) continue;
                    c.tile_threads = (uint32_t)nt;
                    c.tile_points = (uint32_t)t;
                    build_tiles(c);
                    const double ms = krylov_time_launch(c, precision, reps);
                    if (ms < best) { best = ms; best_tile = t; best_threads = nt; }
                }
            }
            c.tile_points = (uint32_t)best_tile;
            c.tile_threads = (uint32_t)best_threads;
            build_tiles(c);
            // staged launches of 512-thread tiles (fiber-only layouts): the
            // current precision's blocks staged or not, whichever launch is
            // shorter (the results are the same bit for bit); the other
            // precision keeps the default (FP32 blocks staged)
            if (precision != CYMG_PRECOND_NONE && c.nsol == 0 && tile_threads_of(c) == 512) {
                const int keep = c.block_stage;
                c.block_stage = 0;
                const double plain = krylov_time_launch(c, precision, reps);
                c.block_stage = 2;
                const double staged = krylov_time_launch(c, precision, reps);
                if (precision == CYMG_PRECOND_FP64) c.block_stage = staged < plain ? 2 : 1;
                else c.block_stage = staged < plain ? (keep == 2 ? 2 : 1) : 0;
                best = std::min(plain, staged);
            }
        }
        c.last_iters[0] = c.last_iters[1] = c.last_iters[2] = 0;
        if (tile_points) *tile_points = best_tile;
        if (block_inverse) *block_inverse = best_method;
        if (launch_ms) *launch_ms = best < 1e300 ? best : 0.0;
    });
}

int cymg_enable_timing(cymg_context* ctx, int enable)
{
    return guarded(ctx, [&](Context& c) { c.timing = enable != 0; });
}

int cymg_time_kernel(cymg_context* ctx, int which, int reps, double* avg_ms)
{
    return guarded(ctx, [&](Context& c) {
        if (!avg_ms || reps <= 0 || which < 0 || which > 7) throw Error(CYMG_ERR_INVALID, "invalid arguments");
        if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede cymg_time_kernel");
        c.solved = false;   // the timed launches run on the solver's vectors
        if (which == 6) {
            if (!c.pc_ready) throw Error(CYMG_ERR_INVALID, "no preconditioner computed");
            *avg_ms = krylov_time_launch(c, c.pc_precision, reps);
            return;
        }
        if (which == 7) {
            *avg_ms = krylov_time_launch(c, CYMG_PRECOND_NONE, reps);
            return;
        }
        if (which == 1 && !c.pc_ready) throw Error(CYMG_ERR_INVALID, "no preconditioner computed");
        const size_t n = c.n();
        c.k_v.resize(n); c.k_t.resize(n);
        CYMG_HIP(hipMemcpyAsync(c.k_v.get(), c.d_rhs.get(), n * 8, hipMemcpyDeviceToDevice, c.stream));
        DevBuf<double> tmp;
        if (which == 4) tmp.resize(c.blk_total);
        const int method = c.pc_method;
        auto once = [&]() {
            if (which == 0) launch_matvec(c, c.k_v.get(), c.k_t.get());
            else if (which == 1) launch_precond_apply(c, c.k_v.get(), c.k_t.get());
            else if (which == 4) launch_precond_build(c, CYMG_PRECOND_FP64, tmp.get());
            else if (which == 5) {
                c.pc_method = 1;
                launch_precond_build(c, CYMG_PRECOND_FP64);
                c.pc_method = method;
            } else launch_precond_build(c, which == 2 ? CYMG_PRECOND_FP64 : CYMG_PRECOND_FP32);
        };
        once();
        CYMG_HIP(hipEventRecord(c.ev[8], c.stream));
        for (int i = 0; i < reps; ++i) once();
        CYMG_HIP(hipEventRecord(c.ev[9], c.stream));
        CYMG_HIP(hipEventSynchronize(c.ev[9]));
        float ms = 0;
        CYMG_HIP(hipEventElapsedTime(&ms, c.ev[8], c.ev[9]));
        *avg_ms = ms / reps;
    });
}

int cymg_guess_init(cymg_context* ctx, uint64_t seed)
{
    return guarded(ctx, [&](Context& c) { mlp_init(c, seed); });
}

int cymg_guess_weights(cymg_context* ctx, int set, float* w)
{
    return guarded(ctx, [&](Context& c) {
        if (!w) throw Error(CYMG_ERR_INVALID, "null weights");
        if (set) mlp_set_weights(c, w);
        else {
            if (!c.mlp.ready) throw Error(CYMG_ERR_INVALID, "no weights loaded");
            mlp_get_weights(c, w);
        }
    });
}

int cymg_guess_record(cymg_context* ctx, uint32_t capacity)
{
    return guarded(ctx, [&](Context& c) { mlp_set_record(c, capacity); });
}

int cymg_guess_samples(cymg_context* ctx, uint32_t* count)
{
    return guarded(ctx, [&](Context& c) {
        if (!count) throw Error(CYMG_ERR_INVALID, "null output");
        *count = c.mlp.ds_count;
    });
}

int cymg_guess_train(cymg_context* ctx, int steps, int batch, double lr, uint64_t seed, double* loss)
{
    return guarded(ctx, [&](Context& c) {
        if (steps < 0 || !(lr > 0.0)) throw Error(CYMG_ERR_INVALID, "invalid training arguments");
        if (!c.mlp.ready) throw Error(CYMG_ERR_INVALID, "no weights loaded");
        const double l = mlp_train(c, steps, batch, lr, seed);
        if (loss) *loss = l;
    });
}

int cymg_guess_eval(cymg_context* ctx, int batch, uint64_t seed, double* loss)
{
    return guarded(ctx, [&](Context& c) {
        if (!loss) throw Error(CYMG_ERR_INVALID, "null argument");
        if (!c.mlp.ready) throw Error(CYMG_ERR_INVALID, "no weights loaded");
        *loss = mlp_eval(c, batch, seed);
    });
}

int cymg_guess_check(cymg_context* ctx, const float* X, const float* T, int batch, double* loss, float* grads)
{
    return guarded(ctx, [&](Context& c) {
        if (!X || !T || !grads) throw Error(CYMG_ERR_INVALID, "null array");
        if (!c.mlp.ready) throw Error(CYMG_ERR_INVALID, "no weights loaded");
        const double l = mlp_check(c, X, T, batch, grads);
        if (loss) *loss = l;
    });
}

int cymg_guess_features(cymg_context* ctx, float* X, float* scale)
{
    return guarded(ctx, [&](Context& c) {
        if (!X || !scale) throw Error(CYMG_ERR_INVALID, "null output");
        if (!c.mlp.step_ready) throw Error(CYMG_ERR_INVALID, "no step features computed");
        mlp_features(c, X, scale);
    });
}

int cymg_guess_predict(cymg_context* ctx, int precondition, double* x0, double* base)
{
    return guarded(ctx, [&](Context& c) {
        if (!x0) throw Error(CYMG_ERR_INVALID, "null output");
        if (!c.setup_done) throw Error(CYMG_ERR_INVALID, "cymg_setup must precede cymg_guess_predict");
        if (precondition < 0 || precondition > 2) throw Error(CYMG_ERR_INVALID, "precondition must be 0, 1 or 2");
        if (!c.mlp.ready) throw Error(CYMG_ERR_UNSUPPORTED, "learned guess not trained or loaded");
        if (precondition) launch_precond_build(c, precondition);
        mlp_prepare_step(c, precondition);
        const double* g = mlp_guess(c);
        d2h(c, x0, g, c.n() * 8);
        if (base) d2h(c, base, c.mlp.y.get(), c.n() * 8);
    });
}

int cymg_get_timings(cymg_context* ctx, cymg_timings* out)
{
    return guarded(ctx, [&](Context& c) {
        if (!out) throw Error(CYMG_ERR_INVALID, "null output");
        *out = c.tm;
    });
}

}  // extern "C"
