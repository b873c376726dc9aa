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
//   KMCCntrl   the kmc{} keys (docs/docs/source/ACDCUserGuide.rst:150-180;
//              member names as read in kmc/KineticMonteCarlo.cpp)
//   Basis::checkBasin  Common/Basis.cpp:731-781; inputBasis, dumpBasis
//              (:356-501, :600-649)
//   Event      kmc/Event.cpp (rate, the packet constructor, setPositions,
//              getDisplacedAtoms)
//   KineticMonteCarlo  kmc/KineticMonteCarlo.cpp: checkForNewBasin,
//              calculateTime, the event choice of runKMC, clearEventList,
//              updateExcludedList, writeEventList
// The MPIComms/Species/Basis/JobControl arguments of the reference become the
// plain arrays they carry (positions SoA, lattice, labels, charges, frozen).
#pragma once

#include <cmath>
#include <fstream>
#include <random>
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
        p.minTimeStep = minTimeStep;
        p.N_delaySteps = N_delaySteps;
        p.N_maxUphill = N_maxUphill;
        p.initialDelay = initialDelay ? 1 : 0;
        p.verletIntegration = verletIntegration ? 1 : 0;
        // Relax::RelaxStructure (optim/Relax.cpp:25-40): "fire", "fire2", or
        // anything else for a single force and energy evaluation
        p.method = method == "fire" ? AKMC_RELAX_FIRE : (method == "fire2" ? AKMC_RELAX_FIRE2 : AKMC_RELAX_NONE);
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

    // positions as one SoA array [3][n]
    std::vector<double> soa() const
    {
        std::vector<double> p(posX);
        p.insert(p.end(), posY.begin(), posY.end());
        p.insert(p.end(), posZ.begin(), posZ.end());
        return p;
    }
    // Basis::inputBasis (Common/Basis.cpp:356-501) from a file; species
    // gives the label of each symbol
    std::string title;
    bool firstWrite = true;
    void inputBasis(const std::string &path, const std::vector<std::string> &species, int &restartIteration,
                    double &restartTime)
    {
        int n = 0, lev = 0;
        check(akmc_io_basis_size(path.c_str(), &n, &lev));
        std::string names = join(species);
        numberOfAtoms = n;
        std::vector<double> p(3 * (size_t)n);
        atmLabel.assign(n, 0);
        frozen.assign(n, 0);
        std::vector<char> t(4096);
        check(akmc_io_read_basis(path.c_str(), names.c_str(), n, atmLabel.data(), p.data(), latVector,
                                 &restartIteration, &restartTime, t.data(), (int)t.size()));
        title = t.data();
        posX.assign(p.begin(), p.begin() + n);
        posY.assign(p.begin() + n, p.begin() + 2 * n);
        posZ.assign(p.begin() + 2 * n, p.end());
    }
    // Basis::dumpBasis (Common/Basis.cpp:600-649), appending to path; the
    // header on this object's first dump, as the reference does
    void dumpBasis(const std::string &path, const std::vector<std::string> &species, double simulationTime,
                   double energy, int iteration, bool append = true)
    {
        std::vector<double> p = soa();
        std::string names = join(species);
        check(akmc_io_dump_basis(path.c_str(), append ? 1 : 0, firstWrite ? 1 : 0, title.c_str(), names.c_str(),
                                 numberOfAtoms, atmLabel.data(), p.data(), latVector, simulationTime, energy,
                                 iteration));
        firstWrite = false;
    }
    static std::string join(const std::vector<std::string> &v)
    {
        std::string s;
        for (const std::string &x : v) s += x + " ";
        return s;
    }

    // Basis::checkBasin: some atom moved by >= detectBasinRadius and none
    // by more than maxBasinRadius
    bool checkBasin(const Basis &basin2, double detectBasinRadius, double maxBasinRadius) const
    {
        std::vector<double> a = soa(), b = basin2.soa();
        int r = akmc_kmc_check_basin(numberOfAtoms, a.data(), b.data(), latVector, detectBasinRadius, maxBasinRadius);
        check(r);
        return r == 1;
    }
};

// KMC control keys; the reference's KMCCntrl.h is missing, so defaults the
// example control file does not set are this library's choice
struct KMCCntrl {
    std::string kmcMethod = "dimer";
    int kmcSteps = 2, kmcImages = 5, recycleType = 1, historySize = 10;
    double kmcTemperature = 1000.0, kmcPreFactor = 1.0e13, kmcMinCap = 0.1, kmcWindow = 50.0, kmcMaxCap = 0.0;
    double basin2Delta = 0.2, kmcBasinRadius = 0.5, kmcMaxRadius = 5.0, recycleRadius = 10.0,
           recyclePercentage = 100.0;
    bool recycle = false, superBasin = false;
    // kmc/kmc.cpp:160
    void setMaxCap() { kmcMaxCap = akmc_kmc_max_cap(kmcMinCap, kmcWindow, kmcTemperature); }
};

// One KMC event: basin-2 and saddle positions of every atom (FP64),
// energies, centre and rate
class Event {
public:
    std::vector<double> bas2PosX, bas2PosY, bas2PosZ, sadPosX, sadPosY, sadPosZ;
    double centre[3] = {0, 0, 0};
    double basin1Energy = 0, basin2Energy = 0, saddleEnergy = 0, actEnergy = 0, rate = 0;

    Event(const Basis &basin2, const Basis &saddle, const double *c, double e1, double e2, double es, double eact,
          double prefactor, double temperature)
        : bas2PosX(basin2.posX), bas2PosY(basin2.posY), bas2PosZ(basin2.posZ), sadPosX(saddle.posX),
          sadPosY(saddle.posY), sadPosZ(saddle.posZ), basin1Energy(e1), basin2Energy(e2), saddleEnergy(es),
          actEnergy(eact), rate(akmc_kmc_rate(eact, prefactor, temperature))
    {
        for (int k = 0; k < 3; ++k) centre[k] = c[k];
    }
    // Event(int* veci, float* vecf, natoms, prefactor, temperature)
    // (kmc/Event.cpp:66-105): an event from the FP32 packet of sendEvent
    std::vector<int> globNumber;
    Event(const int *veci, const float *vecf, int natoms, double prefactor, double temperature)
    {
        std::vector<double> b2(3 * (size_t)natoms), sd(3 * (size_t)natoms);
        double en[4];
        globNumber.assign(natoms, 0);
        check(akmc_kmc_event_unpack(natoms, veci, vecf, prefactor, temperature, globNumber.data(), b2.data(),
                                    sd.data(), en, centre, &rate));
        bas2PosX.assign(b2.begin(), b2.begin() + natoms);
        bas2PosY.assign(b2.begin() + natoms, b2.begin() + 2 * natoms);
        bas2PosZ.assign(b2.begin() + 2 * natoms, b2.end());
        sadPosX.assign(sd.begin(), sd.begin() + natoms);
        sadPosY.assign(sd.begin() + natoms, sd.begin() + 2 * natoms);
        sadPosZ.assign(sd.begin() + 2 * natoms, sd.end());
        basin1Energy = en[0];
        basin2Energy = en[1];
        saddleEnergy = en[2];
        actEnergy = en[3];
    }
    double getRate() const { return rate; }
    double getActivationEnergy() const { return actEnergy; }
    double getBasin2Energy() const { return basin2Energy; }
    const double *getCentre() const { return centre; }
    void setPositions(Basis &bas) const
    {
        bas.posX = bas2PosX;
        bas.posY = bas2PosY;
        bas.posZ = bas2PosZ;
    }
    std::vector<int> getDisplacedAtoms(const Basis &basin, double basinRadius) const
    {
        std::vector<double> p(bas2PosX);
        p.insert(p.end(), bas2PosY.begin(), bas2PosY.end());
        p.insert(p.end(), bas2PosZ.begin(), bas2PosZ.end());
        std::vector<double> b = basin.soa();
        std::vector<int> atoms(basin.numberOfAtoms);
        int m = akmc_kmc_displaced_atoms(basin.numberOfAtoms, p.data(), b.data(), basin.latVector, basinRadius,
                                         atoms.data());
        check(m);
        atoms.resize(m);
        return atoms;
    }
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

// Relax::RelaxStructure(mpi, bas, fld, natoms, const RelaxCntrl&, Energy&, Status&, out):
// dispatches on RelaxCntrl::method ("fire", "fire2", anything else a single
// force and energy evaluation) and verletIntegration (optim/Relax.cpp:19-42)
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

// The host steps of kmc/KineticMonteCarlo.cpp around the device searches.
// Random numbers come from a seeded mt19937_64 through a uniform [0, 1)
// distribution, the generator of kmc/kmc.cpp:27-28,250-253.
class KineticMonteCarlo {
    std::mt19937_64 generator;
    std::uniform_real_distribution<double> dist{0.0, 1.0};

public:
    explicit KineticMonteCarlo(unsigned long long seed = 0) : generator(seed) {}
    double getRandomNumber() { return dist(generator); }

    // calculateTime (:1151-1161): (1/R) ln u, <= 0; runKMC negates it (:173)
    double calculateTime(double totalRate) { return -akmc_kmc_time_increase(totalRate, getRandomNumber()); }

    // the event choice of runKMC (:144-169); -1 when none qualifies
    int selectEvent(const std::vector<Event> &events, double &totalRate)
    {
        std::vector<double> r;
        for (const Event &e : events) r.push_back(e.getRate());
        return akmc_kmc_select((int)r.size(), r.data(), getRandomNumber(), &totalRate);
    }

    // checkForNewBasin (:1626-1713): basin2 starts at the saddle with its free
    // atoms kicked by basin2Delta |D| D (D the minimum-image saddle - basin1
    // displacement of every atom), is relaxed on the device with the
    // saddle's frozen mask, and is a new basin when basin1.checkBasin says so
    bool checkForNewBasin(Field *fld, const KMCCntrl &kmc, const RelaxCntrl &relax, const Basis &basin1,
                          Basis &basin2, const Basis &saddle, Status &relStatus, Energy &basin2Energy)
    {
        const int n = basin1.numberOfAtoms;
        basin2 = saddle;
        const double *L = basin1.latVector;
        // fractional coordinates through the cofactors of the lattice
        const double c0 = L[4] * L[8] - L[5] * L[7], c1 = L[2] * L[7] - L[1] * L[8], c2 = L[1] * L[5] - L[2] * L[4];
        const double c3 = L[5] * L[6] - L[3] * L[8], c4 = L[0] * L[8] - L[2] * L[6], c5 = L[2] * L[3] - L[0] * L[5];
        const double c6 = L[3] * L[7] - L[4] * L[6], c7 = L[1] * L[6] - L[0] * L[7], c8 = L[0] * L[4] - L[1] * L[3];
        const double det = L[0] * c0 + L[1] * c3 + L[2] * c6;
        const double R[9] = {c0 / det, c1 / det, c2 / det, c3 / det, c4 / det, c5 / det, c6 / det, c7 / det, c8 / det};
        std::vector<double> d(3 * (size_t)n);
        double norm = 0.0;
        for (int i = 0; i < n; ++i) {
            const double r[3] = {saddle.posX[i] - basin1.posX[i], saddle.posY[i] - basin1.posY[i],
                                 saddle.posZ[i] - basin1.posZ[i]};
            double f[3];
            for (int j = 0; j < 3; ++j) {
                f[j] = r[0] * R[j] + r[1] * R[3 + j] + r[2] * R[6 + j];
                f[j] -= std::rint(f[j]);
            }
            for (int j = 0; j < 3; ++j) {
                d[j * (size_t)n + i] = f[0] * L[j] + f[1] * L[3 + j] + f[2] * L[6 + j];
                norm += d[j * (size_t)n + i] * d[j * (size_t)n + i];
            }
        }
        norm = std::sqrt(norm) * kmc.basin2Delta;
        for (int i = 0; i < n; ++i) {
            if (!saddle.frozen.empty() && saddle.frozen[i]) continue;
            basin2.posX[i] = saddle.posX[i] + norm * d[i];
            basin2.posY[i] = saddle.posY[i] + norm * d[(size_t)n + i];
            basin2.posZ[i] = saddle.posZ[i] + norm * d[2 * (size_t)n + i];
        }
        Relax rel;
        Status st;
        rel.RelaxStructure(basin2, fld, relax, basin2Energy, st);
        const bool found = basin1.checkBasin(basin2, kmc.kmcBasinRadius, kmc.kmcMaxRadius);
        if (found) relStatus.setStatusSuccess();
        else relStatus.setStatusFailed();
        return found;
    }

    // writeEventList (:1561-1616): events.<iteration> in the working directory
    void writeEventList(const std::vector<std::string> &species, const Basis &basin1,
                        const std::vector<Event> &eventList, double simulationTime, double basin1Energy,
                        int chosenEvent, int iteration)
    {
        const int n = basin1.numberOfAtoms, m = (int)eventList.size();
        std::vector<double> act, cen, es, sad, e2, b2;
        for (const Event &e : eventList) {
            act.push_back(e.actEnergy);
            cen.insert(cen.end(), e.centre, e.centre + 3);
            es.push_back(e.saddleEnergy);
            e2.push_back(e.basin2Energy);
            for (const auto *v : {&e.sadPosX, &e.sadPosY, &e.sadPosZ}) sad.insert(sad.end(), v->begin(), v->end());
            for (const auto *v : {&e.bas2PosX, &e.bas2PosY, &e.bas2PosZ}) b2.insert(b2.end(), v->begin(), v->end());
        }
        std::vector<double> p = basin1.soa();
        std::string names = Basis::join(species), file = "events." + std::to_string(iteration);
        check(akmc_io_write_events(file.c_str(), names.c_str(), n, basin1.atmLabel.data(), p.data(),
                                   basin1.latVector, basin1Energy, m, act.data(), cen.data(), es.data(), sad.data(),
                                   e2.data(), b2.data(), chosenEvent, simulationTime, 1));
    }

    // updateExcludedList (:1291-1313)
    void updateExcludedList(const Basis &basin, const Event &ev, std::vector<int> &excludedList, const KMCCntrl &kmc)
    {
        std::vector<int> atoms = ev.getDisplacedAtoms(basin, kmc.kmcBasinRadius);
        check(akmc_kmc_update_excluded((int)excludedList.size(), excludedList.data(), (int)atoms.size(),
                                       atoms.data(), kmc.historySize));
    }

    // clearEventList (:1163-1265); with superBasin the reference compares the
    // excluded list's entries with the displaced atoms' numbers (:1228-1234)
    void clearEventList(std::vector<Event> &events, const Basis &basin, int chosen, const KMCCntrl &kmc,
                        const std::vector<int> &excludedList)
    {
        const int n = (int)events.size();
        std::vector<double> cen, u(n, 0.0);
        std::vector<int> blocked(n, 0), keep(n, 0);
        for (const Event &e : events) cen.insert(cen.end(), e.centre, e.centre + 3);
        if (kmc.recycle && kmc.recycleType == 2)
            for (int i = 0; i < n; ++i)
                if (i != chosen) u[i] = getRandomNumber();
        if (kmc.superBasin)
            for (int i = 0; i < n; ++i)
                for (int a : events[i].getDisplacedAtoms(basin, kmc.kmcBasinRadius))
                    for (int ex : excludedList)
                        if (ex == a) blocked[i] = 1;
        check(akmc_kmc_recycle(n, cen.data(), chosen, basin.latVector, kmc.recycleRadius, kmc.recycle ? 1 : 0,
                               kmc.recycleType, kmc.recyclePercentage, u.data(), blocked.data(), keep.data()));
        std::vector<Event> kept;
        for (int i = 0; i < n; ++i)
            if (keep[i]) kept.push_back(events[i]);
        events.swap(kept);
    }
};

}  // namespace akmc
