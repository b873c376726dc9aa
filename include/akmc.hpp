// akmc.hpp -- C++ host classes over the C ABI of libakmc_hip.so.
//
// They keep the reference's plugin and control APIs for the force and
// saddle-search path so a caller of the reference changes its includes, not
// its call sites:
//   Field      potentials/metal/Field.h:137-217 and potentials/pair/Field.h:157-234
//              (readPotential, setup, calculateForces, calculateEnergy)
//   Energy     potentials/metal/Energy.h:13-108 (component fields, getTotalEnergy)
//   Status     Common/Status.h:5-54
//   RelaxCntrl Common/RelaxCntrl.h:13-56, ARTCntrl Common/ArtCntrl.h:12-88,
//   DimerCntrl Common/DimerCntrl.h:16-94 (field names and defaults)
//   Relax      optim/Relax.h:116-117 (RelaxStructure)
//   ART, Dimer optim/ART.h:136-138, optim/Dimer.h:68-70 (findSaddlePoint)
// The MPIComms/Species/Basis/JobControl arguments of the reference become the
// plain arrays they carry (positions SoA, lattice, labels, charges, frozen).
#pragma once

#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "akmc.h"

namespace akmc {

inline void check(int rc)
{
    if (rc < 0) throw std::runtime_error(std::string("akmc: ") + akmc_last_error());
}

struct Energy {
    double totalEnergy = 0.0, rcpEnergy = 0.0, realEnergy = 0.0, manyEnergy = 0.0, vdwEnergy = 0.0,
           extEnergy = 0.0;
    double getTotalEnergy() const { return rcpEnergy + realEnergy + vdwEnergy + manyEnergy + extEnergy; }
    void set(const akmc_energy &e)
    {
        totalEnergy = e.total;
        rcpEnergy = e.rcp;
        realEnergy = e.real;
        manyEnergy = e.many;
        vdwEnergy = e.vdw;
        extEnergy = e.ext;
    }
};

class Status {
    int simStatus = AKMC_FAILED;

public:
    int getStatus() const { return simStatus; }
    void setStatusFailed() { simStatus = AKMC_FAILED; }
    void setStatusSuccess() { simStatus = AKMC_SUCCESS; }
    void setStatusInterrupt() { simStatus = AKMC_INTERRUPTED; }
    void setStatusTerminated() { simStatus = AKMC_TERMINATED; }
    void set(int s) { simStatus = s; }
};

struct RelaxCntrl {
    int maxIter = 3000, print = 1;
    double normTol = 1.0e-5, forceTol = 1.0e-4;
    int N_delaySteps = 20, N_maxUphill = 50, N_min = 5;
    double initTimeStep = 0.0005, alphaDecrease = 0.99, stepIncrease = 1.1, stepDecrease = 0.5, alphaStart = 0.1,
           maxTimeStep = 0.00075, minTimeStep = 0.0001;
    bool initialDelay = true, debug = false;
    std::string method = "fire";
    bool verletIntegration = false;
    akmc_relax_params c_params() const
    {
        akmc_relax_params p;
        akmc_relax_defaults(&p);
        p.maxIter = maxIter;
        p.normTol = normTol;
        p.forceTol = forceTol;
        p.N_min = N_min;
        p.initTimeStep = initTimeStep;
        p.alphaDecrease = alphaDecrease;
        p.stepIncrease = stepIncrease;
        p.stepDecrease = stepDecrease;
        p.alphaStart = alphaStart;
        p.maxTimeStep = maxTimeStep;
        return p;
    }
};

struct ARTCntrl {
    int numLanczosVectors = 20, maxActivationSteps = 10;
    double deltaX = 0.01, maxDeltaX = 0.2, initialDisplacement = 0.01, maxEigenvalue = 0.0, eigValTol = 0.001,
           parDamp = 0.5;
    bool goPeprepedicular = false, debug = true;
    int maxParIter = 1000, maxPerpIter = 20, print = 1;
    double normTol = 1.0e-4, forceTol = 1.0e-3, maxStep = 0.01;
    std::string minMethod = "fire";
    int N_min = 5;
    double initTimeStep = 0.005, alphaDecrease = 0.99, stepIncrease = 1.1, stepDecrease = 0.5, alphaStart = 0.1,
           maxTimeStep = 0.01, initTimeStepPerp = 0.005, maxTimeStepPerp = 0.01;
    akmc_art_params c_params() const
    {
        akmc_art_params p;
        akmc_art_defaults(&p);
        p.numLanczosVectors = numLanczosVectors;
        p.deltaX = deltaX;
        p.maxEigenvalue = maxEigenvalue;
        p.eigValTol = eigValTol;
        p.parDamp = parDamp;
        p.maxParIter = maxParIter;
        p.normTol = normTol;
        p.forceTol = forceTol;
        p.N_min = N_min;
        p.initTimeStep = initTimeStep;
        p.alphaDecrease = alphaDecrease;
        p.stepIncrease = stepIncrease;
        p.stepDecrease = stepDecrease;
        p.alphaStart = alphaStart;
        p.maxTimeStep = maxTimeStep;
        return p;
    }
};

struct DimerCntrl {
    double deltaTheta = 0.005, dimerDisplacement = 0.01, lineDisplacement = 0.001, maxDeltaX = 0.5;
    bool debug = true;
    int maxRotations = 20, regionStyle = 1;
    double regionRadius = 5.0;
    int maxParIter = 1000, print = 1;  // 10 in the reference, whose translation loop never runs
    double maxEigenvalue = 0.0, normTol = 0.1, forceTol = 1.0e-2, energyTol = 1.0e-4;
    int N_min = 5;
    double alphaDecrease = 0.99, stepIncrease = 1.1, stepDecrease = 0.5, alphaStart = 0.1;
    // translation time steps and per-atom step cap of this build's dimer
    double initTimeStep = 0.005, maxTimeStep = 0.01, maxStep = 0.1;
    akmc_dimer_params c_params() const
    {
        akmc_dimer_params p;
        akmc_dimer_defaults(&p);
        p.deltaTheta = deltaTheta;
        p.dimerDisplacement = dimerDisplacement;
        p.maxDeltaX = maxDeltaX;
        p.maxRotations = maxRotations;
        p.maxParIter = maxParIter;
        p.maxEigenvalue = maxEigenvalue;
        p.normTol = normTol;
        p.forceTol = forceTol;
        p.N_min = N_min;
        p.alphaDecrease = alphaDecrease;
        p.stepIncrease = stepIncrease;
        p.stepDecrease = stepDecrease;
        p.alphaStart = alphaStart;
        p.initTimeStep = initTimeStep;
        p.maxTimeStep = maxTimeStep;
        p.maxStep = maxStep;
        return p;
    }
};

// The force-field plugin (one class for both the many-body and the ionic
// field; the kind follows from the potentials text).
class Field {
    akmc_field_t h_ = nullptr;

public:
    Field() = default;
    ~Field()
    {
        if (h_) akmc_field_destroy(h_);
    }
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;

    // readPotential(ifstream&, ofstream&, Species&)
    void readPotential(std::istream &in)
    {
        std::stringstream ss;
        ss << in.rdbuf();
        readPotentialText(ss.str());
    }
    void readPotentialText(const std::string &text)
    {
        if (h_) akmc_field_destroy(h_);
        h_ = nullptr;
        check(akmc_field_create(text.c_str(), &h_));
    }
    // setup(mpi, lat[9], rcp[9], minDim, volume, maxAtoms, spec, out)
    void setup(const double *latVector, int maxAtoms) { check(akmc_field_setup(h_, latVector, maxAtoms)); }
    // calculateForces(mpi, posX, posY, posZ, fX, fY, fZ, stress, lat, rcp,
    //                 charge, atmLabel, frozen, numAtoms, Energy&)
    void calculateForces(double *posX, double *posY, double *posZ, double *forceX, double *forceY, double *forceZ,
                         double *stress, const double *latVector, const double *atomCharge, const int *atmLabel,
                         const int *frozen, int numAtoms, Energy &eng)
    {
        akmc_energy e;
        check(akmc_field_forces(h_, posX, posY, posZ, forceX, forceY, forceZ, stress, latVector, atomCharge,
                                atmLabel, frozen, numAtoms, &e));
        eng.set(e);
    }
    void calculateEnergy(double *posX, double *posY, double *posZ, const double *latVector, const double *atomCharge,
                         const int *atmLabel, int numAtoms, Energy &eng)
    {
        akmc_energy e;
        check(akmc_field_energy(h_, posX, posY, posZ, latVector, atomCharge, atmLabel, numAtoms, &e));
        eng.set(e);
    }
    void setEwald(double eta, int gx, int gy, int gz) { check(akmc_field_set_ewald(h_, eta, gx, gy, gz)); }
    int kind() const { return akmc_field_kind(h_); }
    akmc_field_t handle() const { return h_; }
};

// A configuration as the searches need it: the arrays of the reference's
// Basis (Common/Basis.cpp) for one cell.
struct Basis {
    int numberOfAtoms = 0;
    std::vector<double> posX, posY, posZ, charge;
    std::vector<int> atmLabel, frozen;
    double latVector[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

namespace detail {
struct OneWalker {
    akmc_batch_t b = nullptr;
    OneWalker(Field &fld, Basis &bas)
    {
        int n = bas.numberOfAtoms;
        check(akmc_batch_create(fld.handle(), 1, &n, bas.latVector, &b));
        check(akmc_batch_upload(b, bas.posX.data(), bas.posY.data(), bas.posZ.data(), bas.atmLabel.data(),
                                bas.charge.empty() ? nullptr : bas.charge.data(), nullptr));
        if (!bas.frozen.empty()) check(akmc_batch_set_frozen(b, bas.frozen.data()));
    }
    ~OneWalker() { akmc_batch_destroy(b); }
    void read_back(Basis &bas, Energy &e, Status &st, double *curvature = nullptr)
    {
        check(akmc_batch_get_positions(b, bas.posX.data(), bas.posY.data(), bas.posZ.data()));
        akmc_walker_status ws;
        check(akmc_batch_status(b, &ws));
        e.totalEnergy = ws.energy;
        st.set(ws.status);
        if (curvature) *curvature = ws.curvature;
    }
};
}  // namespace detail

// Relax::RelaxStructure(mpi, bas, fld, natoms, const RelaxCntrl&, Energy&, Status&, out)
class Relax {
public:
    void RelaxStructure(Basis &bas, Field *fld, const RelaxCntrl &c, Energy &eng, Status &status)
    {
        detail::OneWalker w(*fld, bas);
        akmc_relax_params p = c.c_params();
        check(akmc_batch_start_relax(w.b, nullptr, &p));
        int steps = 0;
        check(akmc_batch_run(w.b, c.maxIter + 2, 1, &steps));
        w.read_back(bas, eng, status);
    }
};

// findSaddlePoint: the caller supplies the displaced start (saddle holds
// basin1 + displacement) and the unit start direction rVector over all atoms
// ([3][N], frozen components ignored).
class Dimer {
public:
    double curvature = 0.0;
    void findSaddlePoint(Field *fld, Basis &saddle, const std::vector<double> &rVector, const DimerCntrl &c,
                         Status &status, Energy &saddleEnergy)
    {
        detail::OneWalker w(*fld, saddle);
        akmc_dimer_params p = c.c_params();
        check(akmc_batch_start_dimer(w.b, nullptr, nullptr, rVector.data(), nullptr, &p));
        int steps = 0;
        check(akmc_batch_run(w.b, 1 << 30, 1, &steps));
        w.read_back(saddle, saddleEnergy, status, &curvature);
    }
};

class ART {
public:
    double eigenValue = 0.0;
    void findSaddlePoint(Field *fld, Basis &saddle, const std::vector<double> &rVector, const ARTCntrl &c,
                         Status &status, Energy &saddleEnergy)
    {
        if (c.goPeprepedicular) throw std::invalid_argument("the perpendicular ART pre-phase is not offered");
        detail::OneWalker w(*fld, saddle);
        akmc_art_params p = c.c_params();
        check(akmc_batch_start_art(w.b, nullptr, nullptr, rVector.data(), nullptr, &p));
        int steps = 0;
        check(akmc_batch_run(w.b, 1 << 30, 1, &steps));
        w.read_back(saddle, saddleEnergy, status, &eigenValue);
    }
};

}  // namespace akmc
