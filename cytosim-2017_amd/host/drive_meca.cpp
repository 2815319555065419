// Runs implicit steps of a fiber (and aster) scene through the C++ Meca
// interface, the way Simul::solve drives the reference's Meca
// (simul_solve.cc:199-253): clear, add every fiber then every Solid, prepare,
// confinement, couples, asters, solve.
// usage: drive_meca scene.txt out.txt [frame.cmo]
// scene.txt: dim nfib nlinks nsteps nsol naster / time_step kT viscosity
//            tolerance tol_mode radius stiffness / per fiber: np cut rigidity
//            then DIM*np points / per Solid: np timer, np radii, DIM*np points /
//            per link: fiber_a abscissa_a fiber_b abscissa_b stiffness /
//            per aster link: kind s1 s2 fiber abscissa stiffness (s1, s2:
//            matIndex of Solid points; kind 2: abscissa is the interpolation
//            coefficient between s1 and s2) / per step: DIM*nbPts gaussian draws
// out.txt:   per step "iterations residual", then the final points.
// frame.cmo (optional third argument): the final state as a lossless trajectory
//            frame (Simul::writeObjects, simul_file.cc:402-489), read back and
//            checked against the points before it is written.
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "meca.hpp"

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s scene.txt out.txt [frame.cmo]\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1]);
    int dim, nfib, nlinks, nsteps, nsol, nast, tol_mode;
    cymeca::SimulProp prop;
    double radius, stiff;
    in >> dim >> nfib >> nlinks >> nsteps >> nsol >> nast;
    in >> prop.time_step >> prop.kT >> prop.viscosity >> prop.tolerance >> tol_mode >> radius >> stiff;
    std::vector<cymeca::Fiber> fibers(nfib);
    for (auto& f : fibers) {
        unsigned np;
        in >> np >> f.segmentation >> f.rigidity;
        f.points.resize((size_t)dim * np);
        for (double& x : f.points) in >> x;
    }
    std::vector<cymeca::Solid> solids(nsol);
    for (auto& s : solids) {
        unsigned np;
        in >> np >> s.reshapeTimer;
        s.radius.resize(np);
        for (double& r : s.radius) in >> r;
        s.points.resize((size_t)dim * np);
        for (double& x : s.points) in >> x;
        s.shape = s.points;   // Solid::fixShape when built
    }
    std::vector<uint32_t> fa(nlinks), fb(nlinks);
    std::vector<double> aa(nlinks), ab(nlinks), w(nlinks);
    for (int k = 0; k < nlinks; ++k) in >> fa[k] >> aa[k] >> fb[k] >> ab[k] >> w[k];
    std::vector<int> akind(nast);
    std::vector<uint32_t> as1(nast), as2(nast), af(nast);
    std::vector<double> aab(nast), aw(nast);
    for (int k = 0; k < nast; ++k) in >> akind[k] >> as1[k] >> as2[k] >> af[k] >> aab[k] >> aw[k];
    if (!in) {
        std::fprintf(stderr, "bad scene file\n");
        return 2;
    }
    std::ofstream out(argv[2]);
    out << std::setprecision(17);
    try {
        cymeca::Meca meca(dim, 0);
        size_t npts = 0;
        for (auto& f : fibers) npts += f.nbPoints(dim);
        for (auto& s : solids) npts += s.nbPoints();
        std::vector<double> gauss((size_t)dim * npts);
        for (int step = 0; step < nsteps; ++step) {
            for (double& g : gauss) in >> g;
            meca.clear();
            for (auto& f : fibers) meca.add(&f);
            for (auto& s : solids) meca.add(&s);
            meca.prepare(&prop);
            if (radius > 0) meca.confineSphere(radius, stiff);
            for (int k = 0; k < nlinks; ++k)
                meca.interLink(cymeca::interpolate(fibers[fa[k]], dim, aa[k]),
                               cymeca::interpolate(fibers[fb[k]], dim, ab[k]), w[k]);
            // Aster::setInteractions (aster.cc:53-89)
            for (int k = 0; k < nast; ++k) {
                const cymeca::Fiber& f = fibers[af[k]];
                if (akind[k] == 0) meca.interLink(cymeca::PointExact{ as1[k] }, cymeca::PointExact{ f.matIndex }, aw[k]);
                else if (akind[k] == 1)
                    meca.interLink(cymeca::interpolate(f, dim, aab[k]), cymeca::PointExact{ as1[k] }, aw[k]);
                else   // a fiber shorter than the clamp distance: its plus end (aster.cc:77-84)
                    meca.interLink(cymeca::PointInterpolated{ as1[k], as2[k], aab[k] },
                                   cymeca::PointExact{ f.matIndex + f.nbPoints(dim) - 1 }, aw[k]);
            }
            cymg_solve_report rep = meca.solve(&prop, prop.precondition != 0, gauss.data(), tol_mode);
            out << rep.iterations << " " << rep.residual << "\n";
        }
        for (auto& f : fibers)
            for (double x : f.points) out << x << "\n";
        for (auto& s : solids)
            for (double x : s.points) out << x << "\n";
        if (argc > 3) {
            std::vector<cymg_frame_record> recs;
            uint32_t number = 0;
            for (auto& f : fibers) recs.push_back({ ++number, 0, 0, 0, f.segmentation, 0.0 });
            for (size_t k = 0; k < solids.size(); ++k) recs.push_back({ ++number, 0, 1, 0, 0.0, 0.0 });
            meca.setFrameRecords(recs);
            const std::string frame = meca.writeFrame(prop.time_step * nsteps, "drive_meca", CYMG_FRAME_DOUBLE);
            std::vector<cymg_frame_object> objs;
            std::vector<double> pts, rad;
            cymeca::Meca::readFrame(frame, dim, objs, pts, rad);
            bool same = objs.size() == recs.size();
            size_t i = 0;
            for (auto& f : fibers)
                for (double x : f.points) same = same && i < pts.size() && pts[i++] == x;
            for (auto& s : solids)
                for (double x : s.points) same = same && i < pts.size() && pts[i++] == x;
            if (!same || i != pts.size()) {
                std::fprintf(stderr, "the frame read back differs from the points\n");
                return 3;
            }
            std::ofstream(argv[3], std::ios::binary) << frame;
        }
    } catch (const cymeca::Exception& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
