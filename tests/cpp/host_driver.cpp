// Drives the C++ host classes of include/akmc.hpp the way the reference's
// callers do (Field::readPotential/setup/calculateForces,
// Relax::RelaxStructure, Dimer/ART::findSaddlePoint).
//
//   host_driver <action> <potentials file> <config file> <out file>
//   host_driver basis "<species names>" <basis in> <basis out>
//   action: forces | relax | dimer | art | kmc
// config: N, 9 lattice numbers, N lines "x y z type frozen", then for dimer
// and art N lines "mx my mz" (start direction), for kmc N lines of saddle
// positions (the first block is basin 1, its frozen column the saddle's).
// out: "status energy curvature" then N lines of x y z and, for forces,
// N lines of fx fy fz.  kmc: KineticMonteCarlo::checkForNewBasin; status 0
// for a new basin, energy and positions of basin 2.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "akmc.hpp"

int main(int argc, char **argv)
{
    if (argc == 5 && std::string(argv[1]) == "basis") {
        // host_driver basis "<species names>" <basis in> <basis out>:
        // Basis::inputBasis, then Basis::dumpBasis of what was read
        std::vector<std::string> species;
        std::istringstream sp(argv[2]);
        for (std::string w; sp >> w;) species.push_back(w);
        akmc::Basis b;
        int it = 0;
        double t = 0.0;
        try {
            b.inputBasis(argv[3], species, it, t);
            std::ifstream in(argv[3]);
            std::string line;
            std::getline(in, line);
            std::getline(in, line);
            std::getline(in, line);  // the energy sits in the timeStep line
            std::istringstream ls(line);
            std::string w;
            double energy = 0.0;
            for (int k = 0; k < 7 && ls >> w; ++k)
                if (k == 6) energy = std::stod(w);
            b.dumpBasis(argv[4], species, t, energy, it, false);
        } catch (const std::exception &e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
        return 0;
    }
    if (argc != 5) {
        std::fprintf(stderr, "usage: host_driver action potentials config out\n");
        return 2;
    }
    std::string action = argv[1];
    akmc::Field fld;
    std::ifstream pin(argv[2]);
    fld.readPotential(pin);
    std::ifstream cin_(argv[3]);
    akmc::Basis bas;
    cin_ >> bas.numberOfAtoms;
    for (double &v : bas.latVector) cin_ >> v;
    int n = bas.numberOfAtoms;
    bas.posX.resize(n);
    bas.posY.resize(n);
    bas.posZ.resize(n);
    bas.atmLabel.resize(n);
    bas.frozen.resize(n);
    for (int i = 0; i < n; ++i) cin_ >> bas.posX[i] >> bas.posY[i] >> bas.posZ[i] >> bas.atmLabel[i] >> bas.frozen[i];
    std::vector<double> mode;
    if (action == "dimer" || action == "art" || action == "kmc") {
        mode.assign(3 * (size_t)n, 0.0);
        for (int i = 0; i < n; ++i) cin_ >> mode[i] >> mode[n + i] >> mode[2 * (size_t)n + i];
    }
    fld.setup(bas.latVector, n);
    akmc::Energy eng;
    akmc::Status st;
    double curv = 0.0;
    std::vector<double> fx(n), fy(n), fz(n);
    try {
        if (action == "forces") {
            double stress[6];
            fld.calculateForces(bas.posX.data(), bas.posY.data(), bas.posZ.data(), fx.data(), fy.data(), fz.data(),
                                stress, bas.latVector, nullptr, bas.atmLabel.data(), bas.frozen.data(), n, eng);
            st.setStatusSuccess();
        } else if (action == "relax") {
            akmc::Relax rel;
            akmc::RelaxCntrl c;
            rel.RelaxStructure(bas, &fld, c, eng, st);
        } else if (action == "dimer") {
            akmc::Dimer d;
            akmc::DimerCntrl c;
            d.findSaddlePoint(&fld, bas, mode, c, st, eng);
            curv = d.curvature;
        } else if (action == "art") {
            akmc::ART a;
            akmc::ARTCntrl c;  // the reference example's settings (docs/examples/gdc/control:16-23)
            c.initTimeStep = 0.0005;
            c.maxTimeStep = 0.002;
            c.forceTol = 1.0e-2;
            c.parDamp = 0.25;
            c.numLanczosVectors = 25;
            a.findSaddlePoint(&fld, bas, mode, c, st, eng);
            curv = a.eigenValue;
        } else if (action == "kmc") {
            akmc::Basis saddle = bas;
            saddle.posX.assign(mode.begin(), mode.begin() + n);
            saddle.posY.assign(mode.begin() + n, mode.begin() + 2 * n);
            saddle.posZ.assign(mode.begin() + 2 * n, mode.end());
            akmc::Basis basin2;
            akmc::KMCCntrl kc;
            akmc::RelaxCntrl rc;
            akmc::KineticMonteCarlo kmc(7);
            kmc.checkForNewBasin(&fld, kc, rc, bas, basin2, saddle, st, eng);
            bas = basin2;
        } else {
            std::fprintf(stderr, "unknown action %s\n", action.c_str());
            return 2;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    FILE *o = std::fopen(argv[4], "w");
    std::fprintf(o, "%d %.17g %.17g\n", st.getStatus(), eng.totalEnergy, curv);
    for (int i = 0; i < n; ++i) std::fprintf(o, "%.17g %.17g %.17g\n", bas.posX[i], bas.posY[i], bas.posZ[i]);
    if (action == "forces")
        for (int i = 0; i < n; ++i) std::fprintf(o, "%.17g %.17g %.17g\n", fx[i], fy[i], fz[i]);
    std::fclose(o);
    return 0;
}
