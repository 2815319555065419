// C++ host interface of the MI355X implicit step (see meca.hpp).
#include "meca.hpp"

namespace cymeca {

Meca::Meca(int dim, int device, double hr0, double hr1) : dim_(dim), hr0_(hr0), hr1_(hr1)
{
    const int rc = cymg_create(dim, device, &ctx_);
    if (rc) throw Exception(rc, "cymg_create failed (DIM must be 2 or 3 and a GPU visible)");
}

Meca::~Meca()
{
    if (ctx_) cymg_destroy(ctx_);
}

void Meca::check(int rc) const
{
    if (rc) throw Exception(rc, cymg_last_error(ctx_));
}

void Meca::clear()
{
    check(cymg_clear(ctx_));
    objs_.clear();
    solids_.clear();
    nbPts_ = 0;
}

void Meca::add(Fiber* f)
{
    uint32_t mi = 0;
    check(cymg_add_fiber(ctx_, f->nbPoints(dim_), f->segmentation, f->rigidity, &mi));
    f->matIndex = mi;
    objs_.push_back(f);
    nbPts_ += f->nbPoints(dim_);
}

void Meca::add(Solid* s)
{
    uint32_t mi = 0;
    check(cymg_add_solid(ctx_, s->nbPoints(), s->radius.data(), s->shape.data(), s->reshapeTimer, &mi));
    s->matIndex = mi;
    solids_.push_back(s);
    nbPts_ += s->nbPoints();
}

void Meca::prepare(const SimulProp* prop)
{
    cymg_sim_params p{ prop->time_step, prop->kT, prop->viscosity, hr0_, hr1_ };
    check(cymg_set_params(ctx_, &p));
    // putPoints of every Mecable, in matIndex order
    std::vector<double> pts(size());
    for (Fiber* f : objs_)
        std::copy(f->points.begin(), f->points.end(), pts.begin() + (size_t)dim_ * f->matIndex);
    for (Solid* s : solids_)
        std::copy(s->points.begin(), s->points.end(), pts.begin() + (size_t)dim_ * s->matIndex);
    check(cymg_prepare(ctx_, pts.data()));
}

void Meca::interLink(const PointExact& a, const PointExact& b, double w)
{
    check(cymg_inter_link_ee(ctx_, a.mat, b.mat, w));
}

void Meca::interLink(const PointInterpolated& a, const PointExact& b, double w)
{
    check(cymg_inter_link_ie(ctx_, a.mat1, a.mat2, a.coef, b.mat, w));
}

void Meca::interLink(const PointInterpolated& a, const PointInterpolated& b, double w)
{
    check(cymg_inter_link_ii(ctx_, a.mat1, a.mat2, a.coef, b.mat1, b.mat2, b.coef, w));
}

void Meca::interLongLink(const PointExact& a, const PointExact& b, double len, double w)
{
    check(cymg_inter_long_link_ee(ctx_, a.mat, b.mat, len, w));
}

void Meca::interClamp(const PointExact& a, const double g[], double w)
{
    check(cymg_inter_clamp_e(ctx_, a.mat, g, w));
}

void Meca::interClamp(const PointInterpolated& a, const double g[], double w)
{
    check(cymg_inter_clamp_i(ctx_, a.mat1, a.mat2, a.coef, g, w));
}

void Meca::interLongClamp(const PointExact& a, const double center[], double len, double w)
{
    check(cymg_inter_long_clamp(ctx_, a.mat, center, len, w));
}

void Meca::interPlane(const PointExact& a, const double dir[], const double g[], double w)
{
    check(cymg_inter_plane_e(ctx_, a.mat, dir, g, w));
}

void Meca::addPureForce(const PointExact& a, const double force[])
{
    check(cymg_add_pure_force(ctx_, a.mat, force));
}

void Meca::addPureForce(const PointInterpolated& a, const double force[])
{
    check(cymg_add_pure_force_i(ctx_, a.mat1, a.mat2, a.coef, force));
}

void Meca::interPlane(const PointInterpolated& a, const double dir[], const double g[], double w)
{
    check(cymg_inter_plane_i(ctx_, a.mat1, a.mat2, a.coef, dir, g, w));
}

void Meca::interLongLink(const PointInterpolated& a, const PointExact& b, double len, double w)
{
    check(cymg_inter_long_link_ie(ctx_, a.mat1, a.mat2, a.coef, b.mat, len, w));
}

void Meca::interLongLink(const PointInterpolated& a, const PointInterpolated& b, double len, double w)
{
    check(cymg_inter_long_link_ii(ctx_, a.mat1, a.mat2, a.coef, b.mat1, b.mat2, b.coef, len, w));
}

void Meca::interSideLink(const PointInterpolated& a, const PointExact& b, double len, double w)
{
    check(cymg_inter_side_link_ie(ctx_, a.mat1, a.mat2, a.coef, b.mat, len, w));
}

void Meca::interSideLink(const PointInterpolated& a, const PointInterpolated& b, double len, double w)
{
    check(cymg_inter_side_link_ii(ctx_, a.mat1, a.mat2, a.coef, b.mat1, b.mat2, b.coef, len, w));
}

void Meca::interSideSideLink(const PointInterpolated& a, const PointInterpolated& b, double len, double w)
{
    check(cymg_inter_side_side_link(ctx_, a.mat1, a.mat2, a.coef, b.mat1, b.mat2, b.coef, len, w));
}

void Meca::interSlidingLink(const PointInterpolated& a, const PointExact& b, double w)
{
    check(cymg_inter_sliding_link_ie(ctx_, a.mat1, a.mat2, a.coef, b.mat, w));
}

void Meca::interSlidingLink(const PointInterpolated& a, const PointInterpolated& b, double w)
{
    check(cymg_inter_sliding_link_ii(ctx_, a.mat1, a.mat2, a.coef, b.mat1, b.mat2, b.coef, w));
}

void Meca::interSideSlidingLink(const PointInterpolated& a, const PointExact& b, double len, double w)
{
    check(cymg_inter_side_sliding_link_ie(ctx_, a.mat1, a.mat2, a.coef, b.mat, len, w));
}

void Meca::interSideSlidingLink(const PointInterpolated& a, const PointInterpolated& b, double len, double w)
{
    check(cymg_inter_side_sliding_link_ii(ctx_, a.mat1, a.mat2, a.coef, b.mat1, b.mat2, b.coef, len, w));
}

void Meca::interTorque2D(const PointInterpolated& a, const PointInterpolated& b, double cosinus, double sinus,
                         double w)
{
    check(cymg_inter_torque(ctx_, a.mat1, a.mat2, a.coef, b.mat1, b.mat2, b.coef, cosinus, sinus, w));
}

void Meca::interSideClamp(const PointInterpolated& a, const double g[], double len, double w)
{
    check(cymg_inter_side_clamp(ctx_, a.mat1, a.mat2, a.coef, g, len, w));
}

void Meca::interSlidingClamp(const PointInterpolated& a, const double g[], double w)
{
    check(cymg_inter_sliding_clamp(ctx_, a.mat1, a.mat2, a.coef, g, w));
}

void Meca::setSteric(const cymg_steric_params& p, const std::vector<int>& steric, const std::vector<double>& radius,
                     const std::vector<double>& range)
{
    if (radius.size() != steric.size() || range.size() != steric.size())
        throw Exception(CYMG_ERR_INVALID, "one steric flag, radius and range per Mecable");
    check(cymg_set_steric(ctx_, &p, (uint32_t)steric.size(), steric.data(), radius.data(), range.data()));
}

void Meca::stericInteractions()
{
    check(cymg_steric_interactions(ctx_));
}

void Meca::setFrameRecords(const std::vector<cymg_frame_record>& records)
{
    check(cymg_set_frame_records(ctx_, (uint32_t)records.size(), records.data()));
}

std::string Meca::writeFrame(double time, const std::string& date, int real_bytes, const std::string& before,
                             const std::string& after)
{
    cymg_frame_params p = { time, date.c_str(), real_bytes, before.data(), before.size(), after.data(), after.size() };
    size_t n = 0;
    check(cymg_write_frame(ctx_, &p, nullptr, 0, &n));
    std::string out(n, '\0');
    check(cymg_write_frame(ctx_, &p, &out[0], n, &n));
    return out;
}

size_t Meca::readFrame(const std::string& data, int dim, std::vector<cymg_frame_object>& objects,
                       std::vector<double>& points, std::vector<double>& radii, double* time)
{
    uint32_t no = 0;
    size_t np = 0, used = 0;
    // a first pass counts, a second fills
    int rc = cymg_read_frame(data.data(), data.size(), dim, nullptr, 0, &no, nullptr, nullptr, 0, &np, time, &used);
    if (rc) throw Exception(rc, cymg_last_error(nullptr));
    objects.resize(no);
    points.resize((size_t)dim * np);
    radii.resize(np);
    rc = cymg_read_frame(data.data(), data.size(), dim, objects.data(), no, &no, points.data(), radii.data(), np, &np,
                         time, &used);
    if (rc) throw Exception(rc, cymg_last_error(nullptr));
    return used;
}

void Meca::confineSphere(double radius, double stiffness)
{
    check(cymg_confine_sphere(ctx_, radius, stiffness));
}

void Meca::setCouples(const std::vector<uint32_t>& fa, const std::vector<double>& aa,
                      const std::vector<uint32_t>& fb, const std::vector<double>& ab,
                      const std::vector<double>& w)
{
    check(cymg_set_couples(ctx_, (uint32_t)fa.size(), fa.data(), aa.data(), fb.data(), ab.data(), w.data()));
}

void Meca::couplesInteractions()
{
    check(cymg_couples_interactions(ctx_));
}

void Meca::setAsterLinks(const std::vector<int>& kind, const std::vector<uint32_t>& s1,
                         const std::vector<uint32_t>& s2, const std::vector<uint32_t>& fiber,
                         const std::vector<double>& ab, const std::vector<double>& w)
{
    check(cymg_set_aster_links(ctx_, (uint32_t)kind.size(), kind.data(), s1.data(), s2.data(), fiber.data(),
                               ab.data(), w.data()));
}

void Meca::asterInteractions()
{
    check(cymg_aster_interactions(ctx_));
}

void Meca::autotune(int reps)
{
    check(cymg_autotune(ctx_, reps, nullptr, nullptr, nullptr));
}

cymg_solve_report Meca::solve(const SimulProp* prop, bool precondition, const double* gauss, int tol_mode,
                              int precision, int guess)
{
    cymg_solve_params p{};
    p.tolerance = prop->tolerance;
    p.tol_mode = tol_mode;
    p.precondition = precondition ? precision : CYMG_PRECOND_NONE;
    p.max_iter = 0;
    p.guess = guess;
    p.fallback = 1;
    p.gauss_on_device = 0;
    cymg_solve_report rep{};
    check(cymg_solve(ctx_, &p, gauss, &rep));
    getPoints();
    return rep;
}

void Meca::computeForces()
{
    check(cymg_compute_forces(ctx_));
}

void Meca::multiply(const double* X, double* Y) const
{
    check(cymg_multiply(ctx_, X, Y));
}

void Meca::precondition(const double* X, double* Y) const
{
    check(cymg_precondition(ctx_, X, Y));
}

void Meca::getVector(int which, double* out) const
{
    check(cymg_get_vector(ctx_, which, out));
}

void Meca::getPoints()
{
    std::vector<double> pts(size());
    check(cymg_get_points(ctx_, pts.data()));
    for (Fiber* f : objs_)
        std::copy(pts.begin() + (size_t)dim_ * f->matIndex,
                  pts.begin() + (size_t)dim_ * (f->matIndex + f->nbPoints(dim_)), f->points.begin());
    for (Solid* s : solids_)
        std::copy(pts.begin() + (size_t)dim_ * s->matIndex,
                  pts.begin() + (size_t)dim_ * (s->matIndex + s->nbPoints()), s->points.begin());
    // Solid::getPoints advanced each Solid's reshape timer on the device
    if (!solids_.empty()) {
        std::vector<int> t(solids_.size());
        check(cymg_get_solid_timers(ctx_, (uint32_t)t.size(), t.data()));
        for (size_t k = 0; k < t.size(); ++k) solids_[k]->reshapeTimer = t[k];
    }
}

void Meca::dump(const char* directory) const
{
    check(cymg_dump(ctx_, directory));
}

}  // namespace cymeca
