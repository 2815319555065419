// Host-side context of one Meca instance: registration, assembly streams and
// the device arrays of the step.
#pragma once

#include "common.h"
#include "mlp.h"

namespace cymg {

// buffers of the FiberGrid query (grid.hip)
struct GridBufs {
    DevBuf<double> places, oabs, odis;
    DevBuf<uint32_t> excl, count, offset, key, key2, val, val2, start, end, ofib, opt;
    DevBuf<unsigned long long> half;
    DevBuf<int> flag;
    DevBuf<unsigned char> temp;
};

struct Context {
    int dim = 3;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    cymg_sim_params prm{0.005, 0.0042, 1.0, 0.0125, 10.0};

    // registration (Meca::add), host copy
    std::vector<uint32_t> h_np, h_idx, h_seg;
    std::vector<uint64_t> h_blkoff;
    std::vector<double> h_cut, h_rig;
    std::vector<double> h_drag;        // per fiber: hydrodynamic radii r0, r1 (0: the context's) and the
                                       // surface model's cylinder height (0: drag in a volume)
    std::vector<uint32_t> h_ids;       // identity of every Mecable (fibers, then Solids), empty if not given
    // the registration the device points and previous solution belong to
    std::vector<uint32_t> h_prev_ids, h_prev_np;
    bool prev_layout = false;          // a previous registration has been uploaded
    bool pos_complete = false;         // every point of the current layout has its device position
    uint32_t nfib = 0, npts = 0, nseg = 0, largest = 0;
    uint64_t blk_total = 0;
    bool layout_dirty = true;
    uint32_t nsys = 1;
    std::vector<uint32_t> h_sys_begin;   // first fiber of each system, nsys+1 entries (empty: one system)
    std::vector<uint32_t> h_sys_begin_used;   // the same, as uploaded (always nsys+1 entries)
    std::vector<cymg_solve_report> sys_reports;
    // tiles of whole fibers for the two-launch iteration (never across systems)
    std::vector<uint32_t> h_tile_fib, h_tile_sys, h_sys_tile;
    uint32_t ntiles = 0;
    uint32_t tile_points = 0;          // target points per tile, 0: chosen from the layout
    uint32_t tile_threads = 512;       // workgroup of the two-launch iteration (256: fiber-only layouts)

    // Solids, registered after every fiber (simul_solve.cc:204-207)
    uint32_t nsol = 0;
    uint32_t fib_pts = 0;              // points of the fibers; Solid points follow
    std::vector<uint32_t> h_sol_np, h_sol_idx;
    std::vector<double> h_sol_rad, h_sol_shape, h_sol_shape_sqr;   // per point, DIM per point, per Solid
    std::vector<int> h_sol_timer;
    DevBuf<uint32_t> d_sol_np, d_sol_idx;
    DevBuf<double> d_sol_rad, d_sol_shape, d_sol_shape_sqr, d_sol_geo, d_sol_A, d_sol_W;
    DevBuf<int> d_sol_timer, d_sol_useW;

    // per-fiber device arrays
    DevBuf<uint32_t> d_np, d_idx, d_seg, d_fib_sys, d_sys_begin;
    DevBuf<uint32_t> d_pfib;           // fiber of each point
    DevBuf<TileDesc> d_tile_desc;
    std::vector<uint32_t> h_fslot;   // each fiber's first segment lane in its tile's packed layout
    DevBuf<uint32_t> d_fslot;
    DevBuf<uint64_t> d_blkoff;
    DevBuf<double> d_cut, d_rig, d_mob, d_rfrig, d_drag;
    DevBuf<uint32_t> d_remap;          // new point -> point of the previous registration (or kInvalid32)
    DevBuf<double> d_old;              // the previous registration's vector while it is remapped

    // vectors of DIM*nbPts (meca.h:101-106) and per-segment arrays
    DevBuf<double> d_pts, d_pos, d_sol, d_bas, d_rhs, d_for, d_tmp, d_prev_sol, d_gauss;
    DevBuf<double> d_diff, d_jd, d_je, d_jd0, d_je0, d_lag, d_pdiff;
    DevBuf<double> d_pcr;   // PCR coefficients of JJ' ((2 kPcrLevels + 1) per segment)
    DevBuf<int> d_status;              // [0]: projection failures
    int* h_status = nullptr;           // pinned copy of d_status, queued at the end of prepare

    // assembly streams (mB, mC, vBAS) in call order
    HostStream hB, hC, hBas;
    std::vector<EmitterMark> marks;
    std::vector<DeferredCall> h_def;   // position-dependent emitters, in call order
    // steric interactions (simul_solve.cc:139-188): parameters, the items of the
    // steric Mecables for the current layout, and the grid's work space
    bool st_set = false;
    cymg_steric_params st_prm{};
    std::vector<int> h_st_flag;
    std::vector<double> h_st_rad, h_st_rng;
    uint64_t st_gen = ~0ull;           // layout generation the items were made for
    int st_gdim[3] = { 1, 1, 1 };
    double st_inf[3] = { 0, 0, 0 }, st_delta[3] = { 0, 0, 0 };
    uint32_t st_ncells = 0, st_nloc = 0, st_npnt = 0;
    DevBuf<uint32_t> st_lpt, st_lfib, st_lflag, st_ppt, st_pmec;
    DevBuf<double> st_lcut, st_lrad, st_lrng, st_prad;
    DevBuf<uint32_t> st_key, st_key2, st_val, st_val2, st_lstart, st_pstart, st_lrank;
    DevBuf<uint32_t> st_cnt, st_cntC, st_cntS, st_ocnt, st_ocntC, st_ocntS;   // per candidate pair
    DevBuf<uint64_t> st_ntask, st_toff;  // pairs per (cell, region slot) and the first of each
    DevBuf<uint32_t> st_slot;            // the (cell, region slot) of each pair
    DevBuf<unsigned char> st_temp;
    DevBuf<DeferredCall> d_st_calls;
    uint32_t* h_st_tot = nullptr;      // pinned: calls, C entries, vBAS entries
    DevBuf<DeferredCall> d_def;
    // trajectory frames: per-Mecable records, and the packed layout they were made for
    std::vector<cymg_frame_record> h_frame_rec;
    uint64_t frame_rec_gen = 0, frame_made_rec = ~0ull, frame_made_layout = ~0ull;
    int frame_made_rb = 0;
    size_t frame_bytes = 0;              // length of the packed sections
    DevBuf<FrameHead> d_frame_head;
    DevBuf<uint32_t> d_frame_pmap;       // record of each point
    DevBuf<unsigned char> d_frame;
    DevBuf<uint64_t> s_key[3];         // stream inputs (B, C, vBAS) in call order
    DevBuf<double> s_val[3];
    size_t s_total[3] = {0, 0, 0};
    DevBuf<uint64_t> s_key_sorted;
    DevBuf<uint32_t> conf_rank, conf_flag;   // confined points before each point (confinement-only streams)
    DevBuf<uint32_t> s_pos;
    DevBuf<uint64_t> c_key, c_key_sorted;
    DevBuf<uint32_t> c_u, c_u_sorted;
    DevBuf<unsigned char> s_temp;
    DevSym mB, mC, mBas;

    // couples (device emitter)
    uint32_t ncouples = 0;
    // generations of the inputs that the device emitters' keys depend on
    uint64_t layout_gen = 0, couples_gen = 0, asters_gen = 0;
    DevBuf<uint32_t> cpl_fa, cpl_fb;
    DevBuf<double> cpl_aa, cpl_ab, cpl_w;
    DevBuf<uint8_t> cpl_state;          // attached hands and motor hands of each couple

    // aster links (device emitter)
    uint32_t nasters = 0;
    DevBuf<int> ast_kind;
    DevBuf<uint32_t> ast_s1, ast_s2, ast_f;
    DevBuf<double> ast_ab, ast_w;

    // preconditioner
    int pc_precision = 0;              // CYMG_PRECOND_*
    int pc_method = 3;                 // 0 Gauss-Jordan in registers, 1 getrf + getri, 2 Gauss-Jordan in LDS,
                                       // 3 blocked Gauss-Jordan with matrix-core panel updates
    DevBuf<uint32_t> d_gj_list;        // fibers by block-size class
    std::vector<uint32_t> gj_start;    // class boundaries in d_gj_list
    uint64_t gj_gen = ~0ull;           // layout generation the lists were made for
    int gj_bs_lds = 0;                 // largest block of the workgroup-local kernels
    // blocks beyond LDS: work space of the blocked Gauss-Jordan, per block by
    // position in their class (matrix, pivot rows and steps, pivot-row copies)
    uint32_t big_bsmax = 0;
    DevBuf<uint64_t> d_big_sofs;
    DevBuf<uint32_t> d_big_iofs;
    DevBuf<double> d_big_A, d_big_AP;
    DevBuf<int> d_big_piv, d_big_step, d_big_fail;
    DevBuf<double> d_binv;
    DevBuf<float> d_binv32;
    DevBuf<int> d_useblk;

    // Krylov work vectors and scalars
    DevBuf<double> k_r, k_rt, k_p, k_v, k_t, k_ph, k_sh;
    DevBuf<double4> k_g1[2];           // two-launch variant: (shat, w, phat, z) records, by parity
    DevBuf<double2> k_g2;              // two-launch variant: (y, z) records
    int krylov_fused = 1;              // 1: two launches per iteration, 0: four
    DevBuf<double> d_part;             // partial reductions [slot][nfib]
    DevBuf<double> d_tpart;            // partial reductions [slot][tile]
    DevBuf<SysScalars> d_scal;
    SysScalars* h_scal = nullptr;      // pinned mirror
    SysScalars* h_snap = nullptr;      // pinned snapshots of queued iteration chunks (2 x 4096)
    hipEvent_t chunk_ev[2]{};
    int last_iters[3] = {0, 0, 0};     // iterations of the previous solve, per preconditioner
    // captured iteration chunks of the two-launch iteration: [preconditioner][2 or 8 iterations],
    // valid while the launch arguments they were captured with stay the same
    int use_graphs = 1;
    int fp32_apply = 0;
    int block_stage = 1;     // staged launches (copying wavefronts per 512-thread tile): 0 off, 1 FP32 blocks, 2 all               // FP32 blocks applied in FP32 arithmetic (four-launch iteration)
    bool solid_tiles = false;          // the Solids have tiles of their own in the two-launch iteration
    hipGraphExec_t iter_graph[3][2]{};
    std::vector<unsigned char> iter_graph_key[3][2];

    bool prepared = false, setup_done = false, pc_ready = false, have_solution = false;
    bool solved = false;               // d_sol holds the solution of the current setup (cymg_apply needs it)

    // learned initial guess
    MlpState mlp;

    // segments near places (the FiberGrid query)
    GridBufs grid;

    // timing
    bool timing = false;
    cymg_timings tm{};
    hipEvent_t ev[10]{};
    double mv_ms_sum = 0, pc_ms_sum = 0;
    int mv_count = 0, pc_count = 0;

    uint32_t n() const { return (uint32_t)dim * npts; }
    uint32_t nmec() const { return nfib + nsol; }   // Mecables: the stride of per-Mecable partial sums
};

}  // namespace cymg
