// Shared definitions of the MI355X Meca library: device buffers, the context
// layout in HBM, and the launchers each translation unit provides.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cymeca.h"

namespace cymg {

constexpr int kWave = 64;               // CDNA wavefront
constexpr int kMaxFiberPoints = CYMG_MAX_FIBER_POINTS;
constexpr int kShortFiberPoints = 128;  // per-fiber launches of layouts without longer fibers
constexpr uint32_t kInvalid32 = 0xFFFFFFFFu;
constexpr uint64_t kInvalidKey = 0xFFFFFFFFFFFFFFFFull;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CYMG_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::cymg::Error(CYMG_ERR_HIP, std::string(#call) + ": " +               \
                                                  hipGetErrorString(e_));               \
    } while (0)

// grow-only device buffer
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t cap = 0;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (ptr) (void)hipFree(ptr); }
    void resize(size_t count) {
        n = count;
        if (count > cap) {
            if (ptr) CYMG_HIP(hipFree(ptr));
            ptr = nullptr;
            size_t c = count + count / 4 + 64;
            CYMG_HIP(hipMalloc(&ptr, c * sizeof(T)));
            cap = c;
        }
    }
    T* get() const { return ptr; }
    size_t bytes() const { return n * sizeof(T); }
    void swap(DevBuf& o) {
        std::swap(ptr, o.ptr);
        std::swap(cap, o.cap);
        std::swap(n, o.n);
    }
};

// Per-Krylov-system scalar state, one record per independent system; two
// copies alternate between consecutive kernels (read one, write the other).
struct alignas(16) SysScalars {
    double rho1, rho2, alpha, omega, beta;
    double tol;         // residual threshold
    double resid;       // last |r|_inf
    double resid_s;     // |s|_inf after the alpha update (breakdown reporting)
    double bnorm;       // |b|_inf
    double noise;       // min over Mecables of mobility*b (noiseLevel / time_step)
    int    iter, iter_max;
    int    done, flag, pending, stage;
};

// slots of per-fiber partial reductions ([slot][fiber])
// A tile of the two-launch iteration: a run of consecutive whole fibers of one
// system.  Tiles partition the fibers in order and never span two systems.
struct TileDesc {
    uint32_t fb;        // first fiber
    uint32_t nft;       // fibers
    uint32_t P0;        // first point
    uint32_t tp;        // points
    uint32_t sys;       // system
    uint32_t tb, te;    // the tiles of that system
    uint32_t S0;        // first segment
    uint32_t ws;        // lanes of the wavefront-packed segment layout (0: not packed)
};

enum { P_RR = 0, P_RMAX, P_DELTA, P_SMAX, P_TT, P_TS, P_RTS, P_RTT, P_NOISE, P_BNORM, P_SLOTS };

// accumulation stream: (key, delta) pairs in call order
struct HostStream {
    std::vector<uint64_t> key;
    std::vector<double> val;
    void clear() { key.clear(); val.clear(); }
};

// device emitter placed in call order: the host-recorded entries before it are
// those below its recorded positions in each host stream
struct EmitterMark {
    int kind;                     // 1 sphere confinement, 2 couples, 3 aster links, 4 deferred calls,
                                  // 5 steric grid calls (made on the device)
    double a, b;                  // emitter parameters (radius, stiffness)
    size_t posB, posC, posBas;    // host stream sizes when the emitter was called
    uint32_t first = 0, count = 0;        // deferred calls [first, first+count) of Context::h_def
    size_t slots[3] = { 0, 0, 0 };        // their stream entries (B, C, vBAS)
};

// An interaction whose entries depend on the positions at assembly time
// (meca_inter.cc), recorded by the host in call order and evaluated on the
// device from the points when the streams are built: every call owns a fixed
// number of entries in each stream (deferred_slots), written in the order of
// the reference's contributions; entries it does not make keep an invalid key.
enum DeferredKind : uint32_t {
    kDefLongLinkEE = 0,   // interLongLink(PointExact, PointExact)
    kDefLongClamp,        // interLongClamp
    kDefLongLinkIE,       // interLongLink(PointInterpolated, PointExact)
    kDefLongLinkII,       // interLongLink(PointInterpolated, PointInterpolated)
    kDefSideLinkIE,       // interSideLink(PointInterpolated, PointExact)
    kDefSideLinkII,       // interSideLink(PointInterpolated, PointInterpolated)
    kDefSideSideLink,     // interSideSideLink (2D)
    kDefSlidingLinkIE,    // interSlidingLink(PointInterpolated, PointExact)
    kDefSlidingLinkII,    // interSlidingLink(PointInterpolated, PointInterpolated)
    kDefSideSlidingIE,    // interSideSlidingLink(PointInterpolated, PointExact)
    kDefSideSlidingII,    // interSideSlidingLink(PointInterpolated, PointInterpolated)
    kDefTorque,           // interTorque2D
    kDefSideClamp,        // interSideClamp
    kDefSlidingClamp,     // interSlidingClamp
    kDefKinds
};

struct DeferredCall {
    uint32_t kind;
    uint32_t p[4];        // model points: a1 a2 b1 b2 (interpolated), a b (exact)
    uint32_t off[3];      // first entry in the B, C and vBAS ranges of its mark
    double ca, cb;        // interpolation coefficients (PointInterpolated::mCoef)
    double len, w;        // resting length / arm, weight
    double g[3];          // fixed position (clamps, long clamp center); cosinus, sinus (torque)
};

// one Mecable's record of a trajectory frame: its bytes before the points
// (section line, newline, reference, fiber fields, point count) and where it goes
struct FrameHead {
    uint64_t off;         // byte offset of the record in the packed sections
    uint32_t first;       // its first point (matIndex)
    uint32_t np;
    uint32_t hlen;        // bytes in head
    uint32_t solid;       // 1: a radius follows each point
    unsigned char head[64];
};

// entries of one deferred call in the B, C and vBAS streams
__host__ __device__ inline void deferred_slots(uint32_t kind, int D, uint32_t s[3])
{
    const uint32_t sym = D * (D + 1) / 2, off = D * (D - 1) / 2;
    s[0] = s[1] = s[2] = 0;
    switch (kind) {
    case kDefLongLinkEE: s[1] = 3 * D + 4 * off; s[2] = 2 * D; break;
    case kDefLongClamp: s[1] = sym; s[2] = D; break;
    case kDefLongLinkIE: s[1] = 6 * D + 9 * off; s[2] = 3 * D; break;
    case kDefLongLinkII: s[1] = 10 * D + 16 * off; s[2] = 4 * D; break;
    case kDefSideLinkIE:
        if (D == 2) { s[0] = 6; s[1] = 6; } else { s[1] = 45; s[2] = 9; }
        break;
    case kDefSideLinkII:
        if (D == 2) { s[0] = 10; s[1] = 10; } else { s[1] = 78; s[2] = 12; }
        break;
    case kDefSideSideLink: s[0] = 10; s[1] = 12; break;
    case kDefSlidingLinkIE: s[1] = 6 * sym + 3 * off; break;
    case kDefSlidingLinkII: s[1] = D == 2 ? 36 : 78; break;
    case kDefSideSlidingIE:
        if (D == 2) { s[1] = 21; } else { s[1] = 45; s[2] = 9; }
        break;
    case kDefSideSlidingII:
        if (D == 2) { s[1] = 36; } else { s[1] = 78; s[2] = 12; }
        break;
    case kDefTorque: s[1] = 36; s[2] = 72; break;
    case kDefSideClamp:
        if (D == 2) { s[0] = 3; s[1] = 2; s[2] = 4; } else { s[1] = 21; s[2] = 6; }
        break;
    case kDefSlidingClamp: s[1] = 3 * sym + off; s[2] = 2 * D; break;
    default: break;
    }
}

// symmetric sparse matrix on the device, built from a stream
struct DevSym {
    uint32_t size = 0;      // rows
    uint32_t nlow = 0;      // unique lower entries
    uint32_t nnz = 0;       // full CSR entries
    bool used = false;      // Meca::use_mB / use_mC
    DevBuf<uint64_t> low_key;   // sorted unique lower keys (col<<32 | row)
    DevBuf<double> low_val;
    DevBuf<uint32_t> rowptr;    // full symmetric CSR
    DevBuf<uint32_t> col;
    DevBuf<double> val;
    DevBuf<uint32_t> map;       // CSR slot -> lower entry
    DevBuf<uint32_t> nruns;     // device count of runs (unique keys incl. invalid)
    // index structure of the last build, reused while the stream's keys are
    // known to be unchanged: stable sort permutation and runs of equal keys
    DevBuf<uint32_t> perm, runoff, runlen;
    DevBuf<uint64_t> ukey;
    std::vector<uint64_t> sig;  // what the keys were built from (empty: no reuse)
};

struct Context;

// motor hands of the couples (MotorProp::complete, motor_prop.cc:44-66)
struct MotorStep {
    double max_speed_dt, var_speed_dt, min_dabs, max_dabs;
    int limit_speed, hold_end;
};

// launchers (assembly.hip)
void build_streams(Context& c);
// launchers (fiber.hip)
void launch_prepare(Context& c);
void launch_fiber_prepare(Context& c);
void launch_emit_confine(Context& c, uint64_t* keyC, double* valC, uint64_t* keyBas,
                         double* valBas, double radius, double stiffness);
void launch_emit_couples(Context& c, uint64_t* keyB, double* valB);
void launch_step_couples(Context& c, const MotorStep& ms);
void launch_emit_asters(Context& c, uint64_t* keyB, double* valB);
// steric.hip: the steric grid's calls as a device emitter mark
void steric_interactions(Context& c);
// frame.hip: the fiber and solid sections of a trajectory frame, packed from d_pos
void write_frame(Context& c, const cymg_frame_params& p, void* out, size_t capacity, size_t* size);
// the message of the last call of this thread that had no context to hold it
// (a null context, cymg_read_frame); "null context" before any
const char* thread_error();
void set_thread_error(const std::string& msg);
// launchers (emit.hip)
void launch_emit_deferred(Context& c, const EmitterMark& mk, uint64_t* keyB, double* valB, uint64_t* keyC,
                          double* valC, uint64_t* keyBas, double* valBas);
void launch_setup(Context& c, const double* gauss);
void launch_matvec(Context& c, const double* X, double* Y);
void launch_forces(Context& c, const double* X, double* Y, bool with_rigidity);
void launch_apply(Context& c);
void launch_gather_points(Context& c, const uint32_t* map, const double* src, double* dst, uint32_t npts);
void launch_tensions(Context& c);
void launch_block_columns(Context& c, const uint32_t* list, uint32_t count, const uint64_t* off, int by_position,
                          double* M);
// launchers (precond.hip)
void launch_precond_build(Context& c, int precision, double* block_out = nullptr);
void launch_precond_apply(Context& c, const double* X, double* Y);
// launchers (solid.hip)
void launch_solid_prepare(Context& c);
void launch_solid_setup(Context& c, const double* gauss);
void launch_solid_matvec(Context& c, const double* X, double* Y);
void launch_solid_forces(Context& c, const double* X, double* Y);
void launch_solid_pc_build(Context& c);
void launch_solid_pc_apply(Context& c, const double* X, double* Y);
void launch_solid_points(Context& c);
// launchers (krylov.hip)
double krylov_time_launch(Context& c, int precond, int launches);
int tile_threads_of(const Context& c);
int tile_max_points(const Context& c, int threads);
bool tiles_fit(const Context& c);
bool two_launch(const Context& c);
void nearby_segments(Context& c, uint32_t nq, const double* places, const uint32_t* exclude, double range,
                     int inclusive, uint32_t* counts, uint32_t capacity, uint32_t* fiber, uint32_t* point,
                     double* abscissa, double* dist_sqr, uint32_t* total);
void nearby_segments_noalloc(Context& c, uint32_t nq, const double* places, const uint32_t* exclude, double range,
                             int inclusive, uint32_t* counts, uint32_t capacity, uint32_t* fiber, uint32_t* point,
                             double* abscissa, double* dist_sqr, uint32_t* total);
void grid_scratch_init(int device);
void krylov_solve(Context& c, int precond, const double* guess_dev, int max_iter, double tol,
                  int tol_mode, cymg_solve_report* rep);

}  // namespace cymg
