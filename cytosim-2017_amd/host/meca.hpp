// C++ host interface of the MI355X implicit step, shaped like the reference's
// `Meca` (src/sim/meca.h:76-327): Mecables are registered with add(), the
// step is prepare() + the inter*() assembly calls + solve(), and multiply() /
// precondition() expose the operator.  Everything runs on the GPU through the
// C ABI of include/cymeca.h; this class owns one cymg_context.
//
// Mecables here are fibers, described by their points, segmentation and
// rigidity (the RigidFiber state Meca reads through putPoints/prepareMecable),
// and Solids (points, per-point radii, the reference shape fixed at
// construction and the reshape timer).  Points come back through getPoints()
// after the solve, reshaped as FiberNaked::getPoints / Solid::getPoints do.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cymeca.h"

namespace cymeca {

// SimulProp fields the step reads (simul_prop.h)
struct SimulProp {
    double time_step = 0.005;
    double viscosity = 1.0;
    double kT = 0.0042;
    double tolerance = 0.05;   // SimulProp::tolerance
    int precondition = 1;      // SimulProp::precondition
};

// PointExact / PointInterpolated (point_exact.h:10-67, point_interpolated.h:26-137)
struct PointExact {
    uint32_t mat;              // matIndex() of the point
};
struct PointInterpolated {
    uint32_t mat1, mat2;       // matIndex1/2
    double coef;               // pos = (1-coef)*P1 + coef*P2
};

// the reference's Exception("convergence failure in solver") etc.
struct Exception : std::runtime_error {
    int code;
    Exception(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Fiber {
    std::vector<double> points;   // DIM * nbPoints
    double segmentation = 0.5;    // fnCut
    double rigidity = 0.0;        // FiberProp::rigidity
    uint32_t matIndex = 0;        // set by Meca::add
    uint32_t nbPoints(int dim) const { return (uint32_t)(points.size() / dim); }
};

// Solid (solid.h): soRadius per point, soShape fixed by fixShape() when built,
// soReshapeTimer; registered after every fiber
struct Solid {
    std::vector<double> points;   // DIM * nbPoints, current
    std::vector<double> shape;    // DIM * nbPoints, the points at fixShape() time
    std::vector<double> radius;   // per point (0: no drag)
    int reshapeTimer = 0;
    uint32_t matIndex = 0;        // set by Meca::add
    uint32_t nbPoints() const { return (uint32_t)radius.size(); }
};

// FiberNaked::interpolate (fiber_naked.cc:1050-1065), abscissa from the minus end
inline PointInterpolated interpolate(const Fiber& f, int dim, double ab)
{
    const uint32_t np = f.nbPoints(dim);
    if (ab <= 0) return { f.matIndex, f.matIndex + 1, 0.0 };
    const double q = ab / f.segmentation;
    const double n = (double)(uint64_t)q;
    const uint32_t rd = (uint32_t)n;
    if (rd < np - 1) return { f.matIndex + rd, f.matIndex + rd + 1, q - n };
    return { f.matIndex + np - 2, f.matIndex + np - 1, 1.0 };
}

class Meca {
public:
    explicit Meca(int dim, int device = 0, double hydro_radius0 = 0.0125, double hydro_radius1 = 10.0);
    ~Meca();
    Meca(const Meca&) = delete;
    Meca& operator=(const Meca&) = delete;

    void clear();                           // meca.cc:55-60
    void add(Fiber* f);                     // meca.cc:67-79
    void add(Solid* s);                     // meca.cc:67-79, after every fiber
    unsigned nbMecables() const { return (unsigned)(objs_.size() + solids_.size()); }
    unsigned nbPoints() const { return nbPts_; }
    unsigned size() const { return dim_ * nbPts_; }

    void prepare(const SimulProp* prop);    // meca.cc:488-537

    // assembly (meca_inter.cc)
    void interLink(const PointExact& a, const PointExact& b, double weight);
    void interLink(const PointInterpolated& a, const PointExact& b, double weight);
    void interLink(const PointInterpolated& a, const PointInterpolated& b, double weight);
    void interLongLink(const PointExact& a, const PointExact& b, double len, double weight);
    void interClamp(const PointExact& a, const double g[], double weight);
    void interClamp(const PointInterpolated& a, const double g[], double weight);
    void interLongClamp(const PointExact& a, const double center[], double len, double weight);
    void interPlane(const PointExact& a, const double dir[], const double g[], double weight);
    void addPureForce(const PointExact& a, const double force[]);
    void addPureForce(const PointInterpolated& a, const double force[]);
    void interPlane(const PointInterpolated& a, const double dir[], const double g[], double weight);
    // read the positions: evaluated on the device when the matrices are built
    void interLongLink(const PointInterpolated& a, const PointExact& b, double len, double weight);
    void interLongLink(const PointInterpolated& a, const PointInterpolated& b, double len, double weight);
    void interSideLink(const PointInterpolated& a, const PointExact& b, double len, double weight);
    void interSideLink(const PointInterpolated& a, const PointInterpolated& b, double len, double weight);
    void interSideSideLink(const PointInterpolated& a, const PointInterpolated& b, double len, double weight);
    void interSlidingLink(const PointInterpolated& a, const PointExact& b, double weight);
    void interSlidingLink(const PointInterpolated& a, const PointInterpolated& b, double weight);
    void interSideSlidingLink(const PointInterpolated& a, const PointExact& b, double len, double weight);
    void interSideSlidingLink(const PointInterpolated& a, const PointInterpolated& b, double len, double weight);
    void interTorque2D(const PointInterpolated& a, const PointInterpolated& b, double cosinus, double sinus,
                       double weight);
    void interSideClamp(const PointInterpolated& a, const double g[], double len, double weight);
    void interSlidingClamp(const PointInterpolated& a, const double g[], double weight);
    // device-side emitters for the common cases
    void confineSphere(double radius, double stiffness);          // Fiber::setInteractions + SpaceSphere
    // Simul::setStericInteractions (simul_solve.cc:139-188): per Mecable in registration
    // order the steric flag, steric_radius and steric_range
    void setSteric(const cymg_steric_params& p, const std::vector<int>& steric, const std::vector<double>& radius,
                   const std::vector<double>& range);
    void stericInteractions();
    // Simul::writeObjects (simul_file.cc:402-489) in binary mode: per Mecable in registration
    // order its record fields; then the whole frame, the fiber and solid sections packed on the
    // device between the caller's other sections (`before`: space, field; `after`: bead ...
    // organizer), reals as floats (CYMG_FRAME_FLOAT) or lossless doubles (CYMG_FRAME_DOUBLE)
    void setFrameRecords(const std::vector<cymg_frame_record>& records);
    std::string writeFrame(double time, const std::string& date, int real_bytes = CYMG_FRAME_FLOAT,
                           const std::string& before = std::string(), const std::string& after = std::string());
    // the first frame of `data` (cymg_read_frame): records, points (dim per point), radii;
    // returns the bytes up to the frame's end
    static size_t readFrame(const std::string& data, int dim, std::vector<cymg_frame_object>& objects,
                            std::vector<double>& points, std::vector<double>& radii, double* time = nullptr);
    void setCouples(const std::vector<uint32_t>& fa, const std::vector<double>& aa,
                    const std::vector<uint32_t>& fb, const std::vector<double>& ab,
                    const std::vector<double>& stiffness);
    void couplesInteractions();                                   // Couple::setInteractions
    // Aster::setInteractions (aster.cc:53-89): kind 0 interLink(solid point s1, fiber minus end),
    // kind 1 interLink(fiber at abscissa, solid point s1), kind 2 interLink(solid s1-s2 at coef, plus end)
    void setAsterLinks(const std::vector<int>& kind, const std::vector<uint32_t>& s1,
                       const std::vector<uint32_t>& s2, const std::vector<uint32_t>& fiber,
                       const std::vector<double>& abscissa, const std::vector<double>& stiffness);
    void asterInteractions();

    // meca.cc:598-845; gauss: N(0,1) per degree of freedom (nullptr: no Brownian term)
    cymg_solve_report solve(const SimulProp* prop, bool precondition, const double* gauss = nullptr,
                            int tol_mode = CYMG_TOL_REFERENCE, int precond_precision = CYMG_PRECOND_FP64,
                            int guess = CYMG_GUESS_ZERO);
    // the library's autotuner on the current system (after a solve or setup)
    void autotune(int reps = 20);

    void computeForces();                                         // meca.cc:550-580
    void multiply(const double* X, double* Y) const;              // meca.cc:157-181
    void precondition(const double* X, double* Y) const;          // meca.cc:448-466

    // export after solve: vPTS/vFOR and the Mecable points (getPoints)
    void getVector(int which, double* out) const;
    void getPoints();                                             // fills the Mecables' points
    void dump(const char* directory = nullptr) const;             // meca.cc:957-990

    cymg_context* context() const { return ctx_; }

private:
    void check(int rc) const;
    int dim_;
    cymg_context* ctx_ = nullptr;
    std::vector<Fiber*> objs_;
    std::vector<Solid*> solids_;
    unsigned nbPts_ = 0;
    double hr0_, hr1_;
};

}  // namespace cymeca
