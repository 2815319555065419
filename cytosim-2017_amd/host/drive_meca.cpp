// Runs implicit steps of a fiber scene through the C++ Meca interface, the way
// Simul::solve drives the reference's Meca (simul_solve.cc:199-253): clear,
// add every fiber, prepare, confinement, couples, solve.
// usage: drive_meca scene.txt out.txt
// scene.txt: dim nfib nlinks nsteps / time_step kT viscosity tolerance tol_mode
//            radius stiffness / per fiber: np cut rigidity then DIM*np points /
//            per link: fiber_a abscissa_a fiber_b abscissa_b stiffness /
//            per step: DIM*nbPts gaussian draws
// out.txt:   per step "iterations residual", then the final points.
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "meca.hpp"

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s scene.txt out.txt\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1]);
    int dim, nfib, nlinks, nsteps, tol_mode;
    cymeca::SimulProp prop;
    double radius, stiff;
    in >> dim >> nfib >> nlinks >> nsteps;
    in >> prop.time_step >> prop.kT >> prop.viscosity >> prop.tolerance >> tol_mode >> radius >> stiff;
    std::vector<cymeca::Fiber> fibers(nfib);
    for (auto& f : fibers) {
        unsigned np;
        in >> np >> f.segmentation >> f.rigidity;
        f.points.resize((size_t)dim * np);
        for (double& x : f.points) in >> x;
    }
    std::vector<uint32_t> fa(nlinks), fb(nlinks);
    std::vector<double> aa(nlinks), ab(nlinks), w(nlinks);
    for (int k = 0; k < nlinks; ++k) in >> fa[k] >> aa[k] >> fb[k] >> ab[k] >> w[k];
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
        std::vector<double> gauss((size_t)dim * npts);
        for (int step = 0; step < nsteps; ++step) {
            for (double& g : gauss) in >> g;
            meca.clear();
            for (auto& f : fibers) meca.add(&f);
            meca.prepare(&prop);
            if (radius > 0) meca.confineSphere(radius, stiff);
            for (int k = 0; k < nlinks; ++k)
                meca.interLink(cymeca::interpolate(fibers[fa[k]], dim, aa[k]),
                               cymeca::interpolate(fibers[fb[k]], dim, ab[k]), w[k]);
            cymg_solve_report rep = meca.solve(&prop, prop.precondition != 0, gauss.data(), tol_mode);
            out << rep.iterations << " " << rep.residual << "\n";
        }
        for (auto& f : fibers)
            for (double x : f.points) out << x << "\n";
    } catch (const cymeca::Exception& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
