// C ABI of libakmc_hip.so: potentials text, field setup and walker batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/akmc.h"
#include "akmc_internal.h"

namespace {

// e^2/(4 pi eps0) in eV A, the value with which the reference reproduces
// its own example energy (docs/examples/gdc/basis:3)
constexpr double kCTOEV = 14.3997584;
constexpr double kPI = 3.14159265358979323846;

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(AKMC_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

std::vector<std::string> split_ws(const std::string &s)
{
    std::istringstream is(s);
    std::vector<std::string> w;
    std::string t;
    while (is >> t) w.push_back(t);
    return w;
}

std::string lower(std::string s)
{
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

struct SpeciesRec {
    std::string name;
    double mass, charge;
};
struct SCPot {
    std::string A, B;
    double eps, a, n, m, c;
};
// a two-body entry: form (IonForm) and its parameters in the order of the
// potentials text (TwoBody.cpp:98-245)
struct BuckPot {
    std::string A, B;
    double Ab, rho, C;
    int form = 0;
    double par[6] = {0, 0, 0, 0, 0, 0};
};
// a tabulated EAM potential as readLEAM leaves it
// (potentials/metal/MetalPotential.cpp:276-361)
struct EamPot {
    std::string A, B, file;
    int nrho = 0, nr = 0;
    double drho = 0, dr = 0, cut = 0;
    std::vector<double> F, rho, phi;  // embedding, density, pair (the file's r phi divided by r)
};

}  // namespace

struct akmc_field_s {
    int kind = -1;  // 0 many-body, 1 ionic
    double rc = 0.0;
    bool min_image = true;
    bool nbrlist = false;
    double verlet_shell = 1.0;
    int do_ewald = 1;  // pair default (potentials/pair/Field.h:28)
    double madelung_acc = 1.0e-6;
    bool explicit_ewald = false;
    double eta = 0.0;
    int gcell[3] = {0, 0, 0};
    double skin = 0.7;  // best of a 0.45-1.6 A sweep on the batched dimer workload
    double list_density = 0.0;  // atoms per A^3 the neighbour lists are sized for (0: each walker's N / V)
    std::vector<SpeciesRec> species;
    std::vector<SCPot> sc;
    std::vector<BuckPot> buck;
    std::vector<EamPot> eam;
    double2 *eam_dev = nullptr;  // device copy of every EAM table (one allocation)
    double dsf_alpha = 0.0;      // ewald damp: the damping of the shifted Coulomb sum
    double *dsf_dev = nullptr;   // its erfc mesh on the device (energy, force terms)
    FieldDev host{};
    FieldDev *dev = nullptr;
    int table[AKMC_MAXPAIR];
    // cached single-walker batch behind akmc_field_forces
    akmc_batch_t single = nullptr;
    int single_n = -1;
    double single_lat[9] = {0};
    int nk_last = 0;
    int fp64 = 0;
};

struct akmc_batch_s {
    akmc_field_t f = nullptr;
    int W = 0, A = 0;
    std::vector<WalkerDev> hw;
    std::vector<int> hk_n;       // [3][nk]
    std::vector<double> hk_v, hg0, hg1;
    BatchDev d{};
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    std::vector<WalkerOut> hout;
    // LDS staging of the many-body pair kernels is possible when every
    // walker's gather records fit one workgroup's LDS
    bool stage_ok = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long rebuild_base = 0;
    // search state
    SearchDev s{};
    bool search_alloc = false;
    std::vector<WalkerState> hws;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
    SearchDev graph_args{};  // kernel arguments baked into the captured graph
    BatchDev graph_d{};
    int graph_prec = 0;
    bool use_graphs = true;
    // HIP events around the FP32 Sutton-Chen pass-2 launch of every
    // micro-step of akmc_batch_run (kernel timer)
    static constexpr int kTimerSteps = 64;
    bool ktimer = false;
    hipEvent_t kev[2][kTimerSteps] = {};
    double kms = 0.0;
    long long klaunches = 0;
    double *basin = nullptr;  // [3][A] saved search positions
    hipEvent_t evp[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // host copy of the frozen mask and the per-walker free-atom lists
    std::vector<int> hfrozen, hact, hnact;
    // active zones of the search micro-steps (BatchDev.zflag ...): on by
    // default for minimum-image Sutton-Chen fields
    bool zone_on = true;
    // the batch's largest walker has at least kWideLogicAtoms atoms
    bool logic_wide = false;
    int *zflag = nullptr, *zmark = nullptr, *zlist = nullptr, *nz = nullptr;
    std::vector<int> hzflag;
    // kernel variants of the search micro-steps' Sutton-Chen passes:
    // (lanes, unroll, staged) of pass 1, of pass 2, then of the FP64 passes
    // of mixed steps; lanes 0: the default
    int sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // work lists of the search micro-steps' pair kernels and their grids
    int2 *wl_tab[4] = {nullptr, nullptr, nullptr, nullptr};
    int *wl_count = nullptr;
    int wl_grid_staged = 0, wl_grid_plain = 0;
    // pinned status staging: one stream sync per host check of a run
    WalkerState *hws_pin = nullptr;
    int *ov_pin = nullptr;
    char *gat_pin = nullptr;    // pinned / device index lists and packed output
    char *gat_dev = nullptr;    // of akmc_batch_get_walkers_* (gather kernel)
    size_t gat_n = 0;
    int inflight = 0;           // a chunk launched by akmc_batch_run_async awaits its wait
    int inflight_steps = 0;
    void *stage_pin = nullptr;  // pinned host / device staging of walker starts
    void *stage_dev = nullptr;
    size_t stage_n = 0;
    // the fresh states of started walkers (one pinned transfer per start);
    // each pinned buffer is rewritten only after its last transfer is done
    StartItem *st_pin = nullptr, *st_dev = nullptr;
    hipEvent_t stage_ev = nullptr, st_ev = nullptr;
    bool stage_pending = false, st_pending = false;
    // starts queued behind a chunk in flight: their states replace that
    // chunk's status copy when it is waited for
    std::vector<std::pair<int, WalkerState>> started_inflight;
    // reads of ended walkers while a chunk is in flight
    hipStream_t side = nullptr;
    // harvest: each chunk ends by copying the positions of the walkers that
    // ended in it into one of two mapped host buffers ([3][A] each), in turn
    int hv_on = 0, hv_cur = 0, hv_inflight = -1, hv_last = -1;
    int *harvested = nullptr;
    double *hv_pin[2] = {nullptr, nullptr}, *hv_dev[2] = {nullptr, nullptr};
};

namespace {

// readParameters (MetalPotential.cpp:363-386): whole lines of numbers until
// n values are in; numbers past n on the last line are dropped (the
// reference writes them past its array)
bool read_values(std::istream &in, int n, std::vector<double> &out)
{
    out.clear();
    std::string line;
    while ((int)out.size() < n) {
        if (!std::getline(in, line)) return false;
        for (const std::string &w : split_ws(line)) {
            if ((int)out.size() >= n) break;
            out.push_back(std::stod(w));
        }
    }
    return true;
}

// readLEAM (MetalPotential.cpp:276-361): four comment/element lines,
// "Nrho drho Nr dr cutoff", one element line, then F(rho) on the rho mesh,
// rho(r) and r phi(r) on the r mesh; phi_i = (r phi)_i / (i dr), so the
// first entry is the file's value / 0 as in the reference.
int read_lammps_eam(EamPot &p)
{
    std::ifstream in(p.file);
    if (!in) return fail(AKMC_EINVAL, "could not find LAMMPS style EAM file %s", p.file.c_str());
    std::string line;
    try {
        for (int k = 0; k < 4; ++k)
            if (!std::getline(in, line)) throw std::runtime_error("header");
        if (!std::getline(in, line)) throw std::runtime_error("mesh line");
        auto w = split_ws(line);
        p.nrho = std::stoi(w.at(0));
        p.drho = std::stod(w.at(1));
        p.nr = std::stoi(w.at(2));
        p.dr = std::stod(w.at(3));
        p.cut = std::stod(w.at(4));
        if (p.nrho < 2 || p.nr < 2 || !(p.drho > 0) || !(p.dr > 0))
            return fail(AKMC_EINVAL, "EAM file %s: bad mesh line", p.file.c_str());
        if (!std::getline(in, line)) throw std::runtime_error("element line");
        if (!read_values(in, p.nrho, p.F) || !read_values(in, p.nr, p.rho) || !read_values(in, p.nr, p.phi))
            throw std::runtime_error("tables");
    } catch (const std::exception &) {
        return fail(AKMC_EINVAL, "EAM file %s is truncated or malformed", p.file.c_str());
    }
    for (int i = 0; i < p.nr; ++i) p.phi[i] = p.phi[i] / (i * p.dr);
    return AKMC_SUCCESS;
}

int parse_potentials(akmc_field_t f, const char *text)
{
    std::istringstream in(text ? text : "");
    std::string line;
    bool closed = false;
    auto next_line = [&](std::string &l) -> bool {
        while (std::getline(in, l)) {
            if (!split_ws(l).empty()) return true;
        }
        return false;
    };
    while (next_line(line)) {
        auto w = split_ws(lower(line));
        const std::string &k = w[0];
        try {
            if (k == "cutoff") {
                f->rc = std::stod(w.at(1));
            } else if (k == "nbrlist") {
                f->nbrlist = true;
                f->verlet_shell = std::stod(w.at(1));
            } else if (k == "noimage") {
                f->min_image = false;
            } else if (k == "ewald") {
                const std::string &sub = w.at(1);
                if (sub == "precis") {
                    f->do_ewald = 1;
                    f->madelung_acc = std::stod(w.at(2));
                } else if (sub == "damp") {
                    // damped-shifted Coulomb with damping alpha (pair Field.cpp:265-270)
                    f->do_ewald = 2;
                    f->dsf_alpha = std::stod(w.at(2));
                } else if (sub == "none") {
                    f->do_ewald = 0;
                } else {
                    return fail(AKMC_EUNSUPPORTED, "ewald %s is not supported on the device", sub.c_str());
                }
            } else if (k == "species") {
                int n = std::stoi(w.at(1));
                for (int s = 0; s < n; ++s) {
                    std::string sl;
                    if (!next_line(sl)) return fail(AKMC_EINVAL, "species: missing line");
                    auto sw = split_ws(sl);  // species lines keep their case (Species.cpp:45-60)
                    f->species.push_back({sw.at(0), std::stod(sw.at(1)), std::stod(sw.at(2))});
                }
            } else if (k == "many" || k == "manybody") {
                f->kind = 0;
                int n = std::stoi(w.at(1));
                for (int p = 0; p < n; ++p) {
                    std::string kl, pl;
                    if (!next_line(kl) || !next_line(pl)) return fail(AKMC_EINVAL, "manybody: missing lines");
                    std::string key = lower(split_ws(kl).at(0));
                    auto pw = split_ws(pl);
                    if (key == "eam") {
                        // A B file [lammps|dlpoly] (ManyBody.cpp:269-307)
                        EamPot e;
                        e.A = pw.at(0);
                        e.B = pw.at(1);
                        e.file = pw.at(2);
                        std::string fmt = pw.size() > 3 ? pw[3] : "lammps";
                        if (fmt == "dlpoly")
                            return fail(AKMC_EUNSUPPORTED, "the DL_POLY tabulated EAM format is not supported");
                        if (fmt != "lammps") return fail(AKMC_EINVAL, "unrecognised EAM format");
                        const int rc = read_lammps_eam(e);
                        if (rc != AKMC_SUCCESS) return rc;
                        f->eam.push_back(std::move(e));
                    } else if (key == "suttonchen") {
                        // A B eps a n m c (ManyBody.cpp:148-163)
                        f->sc.push_back({pw.at(0), pw.at(1), std::stod(pw.at(2)), std::stod(pw.at(3)),
                                         std::stod(pw.at(4)), std::stod(pw.at(5)), std::stod(pw.at(6))});
                    } else {
                        return fail(AKMC_EUNSUPPORTED, "many-body form '%s' is not supported on the device",
                                    key.c_str());
                    }
                    if (!f->eam.empty() && !f->sc.empty())
                        return fail(AKMC_EINVAL, "many-body potential types can not be mixed");
                }
            } else if (k == "twobody") {
                f->kind = 1;
                int n = std::stoi(w.at(1));
                for (int p = 0; p < n; ++p) {
                    std::string kl, pl;
                    if (!next_line(kl) || !next_line(pl)) return fail(AKMC_EINVAL, "twobody: missing lines");
                    std::string key = lower(split_ws(kl).at(0));
                    // potenKey (TwoBody.cpp:440-470) and the parameter counts of loadPotential
                    int form = -1, npar = 0;
                    if (key == "buck") form = ION_BUCK, npar = 3;
                    else if (key == "morse") form = ION_MORSE, npar = 4;
                    else if (key == "ljones") form = ION_LJONES, npar = 2;
                    else if (key == "bhm") form = ION_BHM, npar = 5;
                    else if (key == "severnsix") form = ION_SEVNSIX, npar = 2;
                    if (form < 0)
                        return fail(AKMC_EUNSUPPORTED, "two-body form '%s' is not supported on the device",
                                    key.c_str());
                    auto pw = split_ws(pl);
                    BuckPot bp{pw.at(0), pw.at(1), 0.0, 1.0, 0.0};
                    bp.form = form;
                    for (int q = 0; q < npar; ++q) bp.par[q] = std::stod(pw.at(2 + q));
                    if (form == ION_BUCK) {  // A B A rho C (TwoBody.cpp:98-112)
                        bp.Ab = bp.par[0];
                        bp.rho = bp.par[1];
                        bp.C = bp.par[2];
                    }
                    f->buck.push_back(bp);
                }
            } else if (k == "close") {
                closed = true;
                break;
            } else if (k == "maxnbrs") {
                // accepted and ignored, as the reference can never match it
            } else {
                return fail(AKMC_EINVAL, "unrecognised key word used in potentials file %s", k.c_str());
            }
        } catch (const std::exception &) {
            return fail(AKMC_EINVAL, "malformed potentials line: %s", line.c_str());
        }
    }
    (void)closed;
    if (f->kind < 0) return fail(AKMC_EINVAL, "potentials text has no manybody or twobody section");
    if (f->species.empty()) return fail(AKMC_EINVAL, "potentials text has no species");
    if ((int)f->species.size() > AKMC_MAXSPEC) return fail(AKMC_ECAPACITY, "more than %d species", AKMC_MAXSPEC);
    if ((int)f->eam.size() > AKMC_MAXEAM) return fail(AKMC_ECAPACITY, "more than %d EAM tables", AKMC_MAXEAM);
    return AKMC_SUCCESS;
}

// species-pair tables: -1 default for many-body (ManyBody.cpp:869), 0 for
// pair potentials (TwoBody.cpp:509)
void build_tables(akmc_field_t f)
{
    int ns = (int)f->species.size();
    FieldDev &h = f->host;
    std::memset(&h, 0, sizeof h);
    h.kind = f->kind;
    h.nspec = ns;
    h.min_image = f->min_image ? 1 : 0;
    h.do_ewald = f->kind == 1 ? f->do_ewald : 0;
    h.rc = f->rc;
    h.rc2 = f->rc * f->rc;
    double rl = f->rc + f->skin;
    h.rlist2 = rl * rl;
    // between rebuilds an atom moves at most skin/2, so every atom that comes
    // within rc of a free atom was within rc + skin/2 of it at the rebuild
    const double rz = f->rc + 0.5 * f->skin + 1.0e-6;
    h.rzone2 = rz * rz;
    h.half_skin2 = 0.25 * f->skin * f->skin;
    // ionic lists put the first shells (r <= 3.5 A) first; the FP32 path
    // evaluates them in FP64
    h.rnear2 = f->kind == 1 ? (3.5 + f->skin) * (3.5 + f->skin) : 0.0;
    h.cfact = kCTOEV;
    for (int i = 0; i < AKMC_MAXPAIR; ++i) f->table[i] = f->kind == 0 ? -1 : 0;
    // EAM: species pair -> table, both orders (calculateEAMEnergyMesh,
    // ManyBody.cpp:937-1000); the device table pointers are set when the
    // tables are uploaded
    h.eam = f->eam.empty() ? 0 : 1;
    for (int i = 0; i < AKMC_MAXPAIR; ++i) h.eam_tab[i] = -1;
    for (int k = 0; k < (int)f->eam.size(); ++k)
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < ns; ++j)
                if (f->species[i].name == f->eam[k].A && f->species[j].name == f->eam[k].B) {
                    h.eam_tab[j * ns + i] = k;
                    h.eam_tab[i * ns + j] = k;
                }
    h.eam_one = 0;
    if (h.eam) {
        h.eam_one = 1;
        for (int i = 0; i < ns * ns; ++i) h.eam_one &= h.eam_tab[i] == h.eam_tab[0] && h.eam_tab[0] >= 0;
        for (int k = 0; k < (int)f->eam.size(); ++k) {
            const EamPot &e = f->eam[k];
            EamTabDev &t = h.eamt[k];
            t.nr = e.nr;
            t.nrho = e.nrho;
            t.dr = e.dr;
            t.drho = e.drho;
            t.cut = e.cut;
        }
    }
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j) {
            int np = f->kind == 0 ? (int)f->sc.size() : (int)f->buck.size();
            for (int k = 0; k < np; ++k) {
                const std::string &A = f->kind == 0 ? f->sc[k].A : f->buck[k].A;
                const std::string &B = f->kind == 0 ? f->sc[k].B : f->buck[k].B;
                if (f->species[i].name == A && f->species[j].name == B) {
                    f->table[j * ns + i] = k;
                    f->table[i * ns + j] = k;
                }
            }
        }
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j) {
            int k = f->table[i * ns + j];
            if (f->kind == 0) {
                SCPairDev &p = h.sc[i * ns + j];
                if (k < 0) continue;
                const SCPot &s = f->sc[k];
                p.on = 1;
                p.d = {s.a, s.n, s.m, s.eps, (s.eps * s.c) * (s.eps * s.c), s.n * s.eps,
                       0.5 * s.m * (s.eps * s.c) * (s.eps * s.c)};
                p.f = {(float)p.d.a, (float)p.d.n, (float)p.d.m, (float)p.d.eps, (float)p.d.rho_coef,
                       (float)p.d.fs_coef, (float)p.d.fm_coef};
                auto as_int = [](double e) { return (e >= 0 && e <= 64 && std::floor(e) == e) ? (int)e : -1; };
                p.n_int = as_int(s.n);
                p.m_int = as_int(s.m);
            } else {
                IonPairDev &p = h.ion[i * ns + j];
                if (k < 0 || k >= (int)f->buck.size()) continue;
                const BuckPot &b = f->buck[k];
                p.on = 1;
                p.form = b.form;
                p.Ad = b.Ab;
                p.inv_rhod = 1.0 / b.rho;
                p.Cd = b.C;
                p.A = (float)p.Ad;
                p.inv_rho = (float)p.inv_rhod;
                p.C = (float)p.Cd;
                for (int q = 0; q < 6; ++q) {
                    p.pd[q] = b.par[q];
                    p.pf[q] = (float)b.par[q];
                }
            }
        }
}

bool orthorhombic(const double *lat)
{
    return lat[1] == 0.0 && lat[2] == 0.0 && lat[3] == 0.0 && lat[5] == 0.0 && lat[6] == 0.0 && lat[7] == 0.0 &&
           lat[0] > 0 && lat[4] > 0 && lat[8] > 0;
}

// eta and reciprocal shells from the Madelung accuracy
// (potentials/pair/Field.cpp:129-132, Ewald.cpp:88-120) for an orthorhombic cell
void ewald_params(const akmc_field_s *f, const double *L, double &eta, int g[3])
{
    if (f->explicit_ewald) {
        eta = f->eta;
        for (int a = 0; a < 3; ++a) g[a] = f->gcell[a];
        return;
    }
    double eps = std::fmin(std::fabs(f->madelung_acc), 0.5);
    double tol = std::sqrt(std::fabs(std::log(eps * f->rc)));
    eta = std::sqrt(std::fabs(std::log(eps * f->rc * tol))) / f->rc;
    double tol1 = std::sqrt(-std::log(eps * f->rc * std::pow(2.0 * tol * eta, 2)));
    for (int a = 0; a < 3; ++a) g[a] = (int)std::rint(0.25 + L[a] * eta * tol1 / kPI);
}

// half-space k set (Ewald.cpp:188-319) for an orthorhombic cell; rcpcut is
// min_a g_a / L_a
void kvectors(const akmc_field_s *f, const double *L, double eta, const int g[3], std::vector<int> &kn,
              std::vector<double> &kv, std::vector<double> &g0, std::vector<double> &g1)
{
    double cut = std::min({g[0] / L[0], g[1] / L[1], g[2] / L[2]});
    double cut2 = cut * cut, vol = L[0] * L[1] * L[2];
    double etasq = eta * eta, gsqfct = 1.0 / (4.0 * etasq);
    double gfct0 = kCTOEV * kPI / (etasq * vol), gfct1 = gsqfct * gfct0;
    for (int nx = 0; nx <= g[0]; ++nx)
        for (int ny = -g[1]; ny <= g[1]; ++ny)
            for (int nz = -g[2]; nz <= g[2]; ++nz) {
                double x = nx / L[0], y = ny / L[1], z = nz / L[2];
                double gsq = x * x + y * y + z * z;
                if (gsq < cut2 && gsq > 1.0e-8) {
                    double kx = 2.0 * kPI * x, ky = 2.0 * kPI * y, kz = 2.0 * kPI * z;
                    double ksq = kx * kx + ky * ky + kz * kz, xg = ksq * gsqfct;
                    double gzz = std::exp(-xg) / xg;
                    kn.push_back(nx);
                    kn.push_back(ny);
                    kn.push_back(nz);
                    kv.push_back(kx);
                    kv.push_back(ky);
                    kv.push_back(kz);
                    g0.push_back(kx < 1.0e-8 ? gzz * gfct0 : gzz * gfct0 * 2.0);
                    g1.push_back(-2.0 * gzz * (1.0 + 1.0 / xg) * gfct1);
                }
            }
    (void)f;
}

int check_cell(const akmc_field_s *f, const double *lat)
{
    if (!orthorhombic(lat)) return fail(AKMC_EUNSUPPORTED, "only orthorhombic cells are supported on the device");
    if (f->rc < 1.0) return fail(AKMC_EINVAL, "short range cutoff is too small : %g", f->rc);
    double mind = std::min({lat[0], lat[4], lat[8]});
    // image-convention checks: metal Field.cpp:53-60, pair Field.cpp:99-106
    double radius = f->kind == 0 ? (f->nbrlist ? f->rc + f->verlet_shell : f->rc) : f->rc + f->verlet_shell;
    if (f->min_image && radius > mind / 2)
        return fail(AKMC_EINVAL, "short range cutoff too large for simulation box: cutoff %g minimum dimension %g",
                    f->rc, mind);
    return AKMC_SUCCESS;
}

template <typename T>
int dalloc(akmc_batch_t b, T **p, size_t n)
{
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void **)p, n * sizeof(T)));
    b->allocs.push_back((void *)*p);
    return AKMC_SUCCESS;
}

#define TRY(expr)                      \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != AKMC_SUCCESS) return rc_; \
    } while (0)

// device copy of the EAM tables (once per field) and their addresses in the
// host FieldDev (again after every build_tables)
int ensure_eam_tables(akmc_field_t f)
{
    if (f->eam.empty()) return AKMC_SUCCESS;
    std::vector<size_t> off_d, off_p, off_e;
    size_t total = 0;
    for (const EamPot &e : f->eam) {
        off_d.push_back(total);
        total += e.nr;
        off_p.push_back(total);
        total += 3 * (size_t)e.nr;
        off_e.push_back(total);
        total += 2 * (size_t)e.nrho;
    }
    if (!f->eam_dev) {
        // value tables as (t_i, t_{i+1} - t_i); derivative tables are the
        // forward differences (MetalPotential.cpp:388-392), last entry 0
        auto deriv = [](const std::vector<double> &t, double h) {
            std::vector<double> d(t.size(), 0.0);
            for (size_t i = 0; i + 1 < t.size(); ++i) d[i] = (t[i + 1] - t[i]) / h;
            return d;
        };
        auto pairs = [](const std::vector<double> &t, size_t i) {
            return make_double2(t[i], i + 1 < t.size() ? t[i + 1] - t[i] : 0.0);
        };
        std::vector<double2> h(total);
        for (size_t k = 0; k < f->eam.size(); ++k) {
            const EamPot &e = f->eam[k];
            const std::vector<double> dphi = deriv(e.phi, e.dr), drho = deriv(e.rho, e.dr), dF = deriv(e.F, e.drho);
            for (int i = 0; i < e.nr; ++i) {
                h[off_d[k] + i] = pairs(e.rho, i);
                h[off_p[k] + 3 * (size_t)i + 0] = pairs(e.phi, i);
                h[off_p[k] + 3 * (size_t)i + 1] = pairs(dphi, i);
                h[off_p[k] + 3 * (size_t)i + 2] = pairs(drho, i);
            }
            for (int i = 0; i < e.nrho; ++i) {
                h[off_e[k] + 2 * (size_t)i + 0] = pairs(e.F, i);
                h[off_e[k] + 2 * (size_t)i + 1] = pairs(dF, i);
            }
        }
        HIPCHK(hipMalloc((void **)&f->eam_dev, total * sizeof(double2)));
        HIPCHK(hipMemcpy(f->eam_dev, h.data(), total * sizeof(double2), hipMemcpyHostToDevice));
    }
    for (size_t k = 0; k < f->eam.size(); ++k) {
        f->host.eamt[k].dens = f->eam_dev + off_d[k];
        f->host.eamt[k].pair = f->eam_dev + off_p[k];
        f->host.eamt[k].emb = f->eam_dev + off_e[k];
    }
    return AKMC_SUCCESS;
}

// the erfc mesh of the damped-shifted Coulomb sum, Field::fieldErfcGen
// (potentials/pair/Field.cpp:19-73): erfc(alpha r)/r by Abramowitz and
// Stegun 7.1.26 and (that + 2 alpha/sqrt(pi) exp(-alpha^2 r^2)) / r^2 at
// r = i rc / (n - 4); the shift terms are the values at rc.  The
// reference's point count (MAXMESH) lives in its missing Constants.h; the
// count here is this library's choice.
constexpr int kDsfMesh = 5000;

int ensure_dsf_mesh(akmc_field_t f)
{
    if (f->kind != 1 || f->do_ewald != 2) return AKMC_SUCCESS;
    const int n = kDsfMesh;
    const double rc = f->rc, a = f->dsf_alpha, del = rc / (n - 4);
    std::vector<double> h(2 * (size_t)n, 0.0);
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027, a5 = 1.061405429,
                 pp = 0.3275911;
    for (int i = 1; i < n; ++i) {
        const double r = (double)i * del, rsq = r * r;
        const double tt = 1.0 / (1.0 + pp * a * r);
        const double ex = std::exp(-std::pow(a * r, 2.0));
        h[i] = tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * ex / r;
        h[n + i] = (h[i] + 2.0 * (a / 1.7724538509055159) * ex) / rsq;
    }
    if (!f->dsf_dev) {
        HIPCHK(hipMalloc((void **)&f->dsf_dev, 2 * (size_t)n * sizeof(double)));
        HIPCHK(hipMemcpy(f->dsf_dev, h.data(), 2 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    }
    FieldDev &d = f->host;
    d.dsf_e = f->dsf_dev;
    d.dsf_f = f->dsf_dev + n;
    d.dsf_n = n;
    d.dsf_del = del;
    d.dsf_rfld0 = h[n - 4];
    d.dsf_rfld1 = rc * h[n + n - 4];
    d.eta = a;
    return AKMC_SUCCESS;
}

int batch_build(akmc_field_t f, int W, const int *natoms, const double *lat, akmc_batch_t b)
{
    b->f = f;
    b->W = W;
    b->hw.resize(W);
    int A = 0, ncells = 0, G1 = 1, maxnk = 0;
    double rl = f->rc + f->skin;
    double max_density = 0.0;
    for (int w = 0; w < W; ++w) {
        const double *lw = lat + 9 * w;
        TRY(check_cell(f, lw));
        if (natoms[w] <= 0 || natoms[w] > AKMC_MAX_WALKER_ATOMS)
            return fail(AKMC_EINVAL, "walker %d: atom count %d out of range", w, natoms[w]);
        WalkerDev &h = b->hw[w];
        std::memset(&h, 0, sizeof h);
        h.off = A;
        h.n = natoms[w];
        A += h.n;
        h.L[0] = lw[0];
        h.L[1] = lw[4];
        h.L[2] = lw[8];
        for (int a = 0; a < 3; ++a) {
            h.invL[a] = 1.0 / h.L[a];
            h.nc[a] = std::min(AKMC_MAX_CELLS_AXIS, std::max(1, (int)std::floor(h.L[a] / rl)));
            h.shells[a] = (int)(rl / h.L[a]) + 1;
            if (h.shells[a] > 7) return fail(AKMC_ECAPACITY, "walker %d: more than 7 image shells", w);
        }
        h.cell_off = ncells;
        ncells += h.nc[0] * h.nc[1] * h.nc[2];
        max_density = std::max(max_density, h.n / (h.L[0] * h.L[1] * h.L[2]));
        max_density = std::max(max_density, f->list_density);
        if (f->kind == 1 && f->do_ewald == 1) {
            double eta;
            int g[3];
            ewald_params(f, h.L, eta, g);
            f->host.eta = eta;
            std::vector<int> kn;
            std::vector<double> kv, g0, g1;
            kvectors(f, h.L, eta, g, kn, kv, g0, g1);
            h.k_off = (int)b->hg0.size();
            h.nk = (int)g0.size();
            h.gmax = std::max({g[0], g[1], g[2]});
            G1 = std::max(G1, h.gmax + 1);
            maxnk = std::max(maxnk, h.nk);
            b->hk_n.insert(b->hk_n.end(), kn.begin(), kn.end());
            b->hk_v.insert(b->hk_v.end(), kv.begin(), kv.end());
            b->hg0.insert(b->hg0.end(), g0.begin(), g0.end());
            b->hg1.insert(b->hg1.end(), g1.begin(), g1.end());
            f->nk_last = h.nk;
        }
    }
    b->A = A;
    int nk = (int)b->hg0.size();
    // neighbour capacity: 1.6x the expected count in the list sphere + 64
    double expect = 4.0 / 3.0 * kPI * rl * rl * rl * max_density;
    int maxnb = ((int)(1.6 * expect) + 64 + 31) / 32 * 32;
    // list offsets are 32-bit row A + i products (24-bit operands)
    if (A >= (1 << 24) || (size_t)(maxnb + AKMC_NB_PAD_ROWS) * (size_t)A >= ((size_t)1 << 32))
        return fail(AKMC_ECAPACITY, "batch of %d atoms with %d list entries per atom is too large", A, maxnb);
    BatchDev &d = b->d;
    std::memset(&d, 0, sizeof d);
    d.A = A;
    d.W = W;
    d.maxnb = maxnb;
    d.lpa = 1;
    d.unroll = 1;
    {
        int nmax = 0;
        for (int w = 0; w < W; ++w) nmax = std::max(nmax, natoms[w]);
        d.nb16 = (f->kind == 0 && f->min_image && nmax <= 65535) ? 1 : 0;
        // above this size a 256-thread workgroup loops 64 times over a
        // walker's atoms per vector operation of its state machine
        constexpr int kWideLogicAtoms = 16384;
        b->logic_wide = nmax >= kWideLogicAtoms;
    }
    d.ncells_total = ncells;
    d.nk_total = nk;
    d.G1 = G1;
    d.maxnk = maxnk;
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&b->ev0));
    HIPCHK(hipEventCreate(&b->ev1));
    WalkerDev *wd;
    int *wk, *tp;
    double *q;
    TRY(dalloc(b, &wd, W));
    TRY(dalloc(b, &wk, A));
    TRY(dalloc(b, &tp, A));
    TRY(dalloc(b, &q, A));
    d.walkers = wd;
    d.wk = wk;
    d.type = tp;
    d.q = q;
    TRY(dalloc(b, &d.p4, A));
    TRY(dalloc(b, &d.x, A));
    TRY(dalloc(b, &d.y, A));
    TRY(dalloc(b, &d.z, A));
    TRY(dalloc(b, &d.xr, A));
    TRY(dalloc(b, &d.yr, A));
    TRY(dalloc(b, &d.zr, A));
    TRY(dalloc(b, &d.fx, A));
    TRY(dalloc(b, &d.fy, A));
    TRY(dalloc(b, &d.fz, A));
    TRY(dalloc(b, &d.nbr, (size_t)(maxnb + AKMC_NB_PAD_ROWS) * A));
    TRY(dalloc(b, &d.nnb, A));
    TRY(dalloc(b, &d.nnear, A));
    TRY(dalloc(b, &d.rebuild, W));
    TRY(dalloc(b, &d.overflow, 1));
    TRY(dalloc(b, &d.cell_count, ncells));
    TRY(dalloc(b, &d.cell_start, ncells));
    TRY(dalloc(b, &d.cell_atoms, A));
    TRY(dalloc(b, &d.rho, A));
    TRY(dalloc(b, &d.inv_sqrt_rho, A));
    TRY(dalloc(b, &d.isr64, A));
    TRY(dalloc(b, &d.e_a, A));
    TRY(dalloc(b, &d.e_b, A));
    TRY(dalloc(b, &d.e_a64, A));
    TRY(dalloc(b, &d.e_b64, A));
    TRY(dalloc(b, &d.e_c, A));
    TRY(dalloc(b, &d.st_atom, (size_t)6 * A));
    TRY(dalloc(b, &d.out, W));
    TRY(dalloc(b, &d.active, W));
    if (nk > 0) {
        int *kn;
        double *kv, *g0, *g1;
        TRY(dalloc(b, &kn, (size_t)3 * nk));
        TRY(dalloc(b, &kv, (size_t)3 * nk));
        TRY(dalloc(b, &g0, nk));
        TRY(dalloc(b, &g1, nk));
        // transpose to [3][nk]
        std::vector<int> tn(3 * (size_t)nk);
        std::vector<double> tv(3 * (size_t)nk);
        for (int k = 0; k < nk; ++k)
            for (int a = 0; a < 3; ++a) {
                tn[(size_t)a * nk + k] = b->hk_n[3 * (size_t)k + a];
                tv[(size_t)a * nk + k] = b->hk_v[3 * (size_t)k + a];
            }
        HIPCHK(hipMemcpy(kn, tn.data(), tn.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(kv, tv.data(), tv.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(g0, b->hg0.data(), nk * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(g1, b->hg1.data(), nk * sizeof(double), hipMemcpyHostToDevice));
        d.kn = kn;
        d.kv = kv;
        d.g0 = g0;
        d.g1 = g1;
        TRY(dalloc(b, &d.ptab, (size_t)A * 3 * G1));
        TRY(dalloc(b, &d.Ck, nk));
        TRY(dalloc(b, &d.Sk, nk));
    }
    std::vector<int> hwk(A);
    for (int w = 0; w < W; ++w)
        for (int i = 0; i < b->hw[w].n; ++i) hwk[b->hw[w].off + i] = w;
    HIPCHK(hipMemcpy(wd, b->hw.data(), W * sizeof(WalkerDev), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(wk, hwk.data(), A * sizeof(int), hipMemcpyHostToDevice));
    {
        // LDS staging: a walker's packed gather records (32 bytes per atom)
        // next to the static pair table of the multi-species kernels
        int nmax = 0, dev = 0, maxlds = 0;
        for (int w = 0; w < W; ++w) nmax = std::max(nmax, b->hw[w].n);
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t bytes = (size_t)nmax * sizeof(double4);
        b->stage_ok = f->kind == 0 && bytes + sizeof(SCPairDev) * AKMC_MAXPAIR <= (size_t)maxlds;
        d.stage_bytes = (int)bytes;
        for (int li = 0; li < 4; ++li) {
            const int apc = AKMC_STAGE_BLOCK >> li;  // atoms per chunk at 1, 2, 4, 8 lanes per atom
            std::vector<int2> ch;
            for (int w = 0; w < W; ++w)
                for (int a0 = 0; a0 < b->hw[w].n; a0 += apc) ch.push_back(make_int2(w, a0));
            int2 *dc;
            TRY(dalloc(b, &dc, ch.size()));
            HIPCHK(hipMemcpy(dc, ch.data(), ch.size() * sizeof(int2), hipMemcpyHostToDevice));
            d.chunk_tab[li] = dc;
            d.nchunk_tab[li] = (int)ch.size();
        }
    }
    {
        std::vector<int> hcw(ncells);
        for (int w = 0; w < W; ++w) {
            const WalkerDev &h = b->hw[w];
            for (int c = 0; c < h.nc[0] * h.nc[1] * h.nc[2]; ++c) hcw[h.cell_off + c] = w;
        }
        int *cw;
        TRY(dalloc(b, &cw, ncells));
        HIPCHK(hipMemcpy(cw, hcw.data(), ncells * sizeof(int), hipMemcpyHostToDevice));
        d.cell_wk = cw;
        TRY(dalloc(b, &d.cell_work, (size_t)ncells + A / 16 + 1));  // list-build items (cell, home chunk)
        TRY(dalloc(b, &d.cell_work_n, 1));
        HIPCHK(hipMemset(d.cell_work_n, 0, sizeof(int)));
    }
    HIPCHK(hipMemset(d.rebuild, 0, W * sizeof(int)));
    {
        std::vector<int> one(W, f->fp64 ? 2 : 1);
        HIPCHK(hipMemcpy(d.active, one.data(), W * sizeof(int), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemset(d.overflow, 0, sizeof(int)));
    HIPCHK(hipMemset(d.e_c, 0, A * sizeof(double)));
    HIPCHK(hipMemset((void *)d.q, 0, A * sizeof(double)));
    // device copy of the field (the eta of the last walker; all walkers of a
    // batch share the field's eta rule)
    if (!f->dev) HIPCHK(hipMalloc(&f->dev, sizeof(FieldDev)));
    TRY(ensure_eam_tables(f));
    TRY(ensure_dsf_mesh(f));
    HIPCHK(hipMemcpy(f->dev, &f->host, sizeof(FieldDev), hipMemcpyHostToDevice));
    d.field = f->dev;
    b->hout.resize(W);
    unsigned long long *cnt = akmc_rebuild_counter_ptr();
    if (!cnt) return fail(AKMC_EHIP, "rebuild counter allocation failed");
    d.rebuild_count = cnt;
    HIPCHK(hipMemcpy(&b->rebuild_base, cnt, sizeof b->rebuild_base, hipMemcpyDeviceToHost));
    return AKMC_SUCCESS;
}

void batch_free(akmc_batch_t b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    for (auto &e : b->evp)
        if (e) (void)hipEventDestroy(e);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (auto &row : b->kev)
        for (auto &e : row)
            if (e) (void)hipEventDestroy(e);
    if (b->hws_pin) (void)hipHostFree(b->hws_pin);
    if (b->ov_pin) (void)hipHostFree(b->ov_pin);
    if (b->gat_pin) (void)hipHostFree(b->gat_pin);
    if (b->gat_dev) (void)hipFree(b->gat_dev);
    if (b->stage_pin) (void)hipHostFree(b->stage_pin);
    if (b->stage_dev) (void)hipFree(b->stage_dev);
    if (b->st_pin) (void)hipHostFree(b->st_pin);
    if (b->stage_ev) (void)hipEventDestroy(b->stage_ev);
    if (b->st_ev) (void)hipEventDestroy(b->st_ev);
    if (b->side) (void)hipStreamDestroy(b->side);
    for (double *p : b->hv_pin)
        if (p) (void)hipHostFree(p);
    delete b;
}

int batch_mode(akmc_batch_t b) { return b->f->fp64 ? 2 : 1; }

int run_forces(akmc_batch_t b, int want_stress)
{
    akmc_launch_forces_impl(b->d, b->f->host, batch_mode(b), want_stress, b->stream);
    HIPCHK(hipGetLastError());
    return AKMC_SUCCESS;
}

// the error of a nonzero overflow word
int overflow_error(akmc_batch_t b, int ov)
{
    if (ov & AKMC_OV_LIST) return fail(AKMC_ECAPACITY, "neighbour list capacity %d exceeded", b->d.maxnb);
    if (ov & AKMC_OV_EAM_RHO) return fail(AKMC_EINVAL, "the rho for embedding is out of bounds of its EAM table");
    return AKMC_SUCCESS;
}

int check_overflow(akmc_batch_t b)
{
    int ov = 0;
    HIPCHK(hipMemcpyAsync(&ov, b->d.overflow, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return overflow_error(b, ov);
}

int fetch_out(akmc_batch_t b)
{
    HIPCHK(hipMemcpyAsync(b->hout.data(), b->d.out, b->W * sizeof(WalkerOut), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}

void to_energy(const WalkerOut &o, akmc_energy *e)
{
    e->total = o.total;
    e->rcp = o.rcp;
    e->real = o.real;
    e->many = o.many;
    e->vdw = o.vdw;
    e->ext = 0.0;
}

// in-place wraps of the caller's arrays, as Field::calculateForces does
// (potentials/metal/Field.cpp:1855-1920)
void host_wrap(bool min_image, const double *lat, int n, double *x, double *y, double *z)
{
    double L[3] = {lat[0], lat[4], lat[8]};
    for (int i = 0; i < n; ++i) {
        double s[3] = {x[i] / L[0], y[i] / L[1], z[i] / L[2]};
        for (int a = 0; a < 3; ++a) {
            if (min_image) {
                s[a] -= std::rint(s[a]);
            } else {
                if (s[a] < 0.0) s[a] += 1.0;
                if (s[a] >= 1.0) s[a] -= 1.0;
            }
        }
        x[i] = s[0] * L[0];
        y[i] = s[1] * L[1];
        z[i] = s[2] * L[2];
    }
}

int ensure_single(akmc_field_t f, int n, const double *lat)
{
    bool same = f->single && f->single_n == n && std::memcmp(f->single_lat, lat, sizeof f->single_lat) == 0;
    if (same) return AKMC_SUCCESS;
    if (f->single) batch_free(f->single);
    f->single = new akmc_batch_s();
    int rc = batch_build(f, 1, &n, lat, f->single);
    if (rc != AKMC_SUCCESS) {
        batch_free(f->single);
        f->single = nullptr;
        return rc;
    }
    f->single_n = n;
    std::memcpy(f->single_lat, lat, sizeof f->single_lat);
    return AKMC_SUCCESS;
}

int upload(akmc_batch_t b, const double *x, const double *y, const double *z, const int *type, const double *q)
{
    size_t A = b->A;
    int ns = (int)b->f->species.size();
    for (size_t i = 0; i < A; ++i)
        if (type[i] < 0 || type[i] >= ns) return fail(AKMC_EINVAL, "atom %zu: species %d out of range", i, type[i]);
    HIPCHK(hipMemcpyAsync(b->d.x, x, A * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d.y, y, A * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d.z, z, A * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync((void *)b->d.type, type, A * sizeof(int), hipMemcpyHostToDevice, b->stream));
    if (q) HIPCHK(hipMemcpyAsync((void *)b->d.q, q, A * sizeof(double), hipMemcpyHostToDevice, b->stream));
    // a fresh configuration always gets fresh lists, and every walker is
    // evaluated by plain force calls
    std::vector<int> ones(b->W, 1), mode(b->W, batch_mode(b));
    HIPCHK(hipMemcpyAsync(b->d.rebuild, ones.data(), b->W * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d.active, mode.data(), b->W * sizeof(int), hipMemcpyHostToDevice, b->stream));
    // a running search's zone does not survive new positions
    if (b->zflag) HIPCHK(hipMemsetAsync(b->zflag, 0, b->W * sizeof(int), b->stream));
    if (b->search_alloc) {
        HIPCHK(hipMemcpyAsync(b->s.X, x, A * sizeof(double), hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipMemcpyAsync(b->s.X + A, y, A * sizeof(double), hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipMemcpyAsync(b->s.X + 2 * A, z, A * sizeof(double), hipMemcpyHostToDevice, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}


// rebuild the free-atom lists of the selected walkers (all when null) from
// the host copy of the frozen mask and upload them
int upload_active(akmc_batch_t b, const int *walkers)
{
    for (int w = 0; w < b->W; ++w) {
        if (walkers && !walkers[w]) continue;
        const int off = b->hw[w].off, n = b->hw[w].n;
        int m = 0;
        for (int i = 0; i < n; ++i)
            if (!b->hfrozen[off + i]) b->hact[off + m++] = off + i;
        b->hnact[w] = m;
        if (m > 0)
            HIPCHK(hipMemcpyAsync((void *)(b->s.act + off), b->hact.data() + off, m * sizeof(int),
                                  hipMemcpyHostToDevice, b->stream));
    }
    HIPCHK(hipMemcpyAsync((void *)b->s.nact, b->hnact.data(), b->W * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}

int search_alloc(akmc_batch_t b)
{
    if (b->search_alloc) return AKMC_SUCCESS;
    size_t A3 = 3 * (size_t)b->A;
    TRY(dalloc(b, &b->s.X, A3));
    TRY(dalloc(b, &b->s.N, A3));
    TRY(dalloc(b, &b->s.V, A3));
    TRY(dalloc(b, &b->s.F0, A3));
    TRY(dalloc(b, &b->s.F1, A3));
    TRY(dalloc(b, &b->s.T, A3));
    int *fr;
    TRY(dalloc(b, &fr, b->A));
    HIPCHK(hipMemset(fr, 0, b->A * sizeof(int)));
    b->s.frozen = fr;
    int *act, *nact;
    TRY(dalloc(b, &act, b->A));
    TRY(dalloc(b, &nact, b->W));
    TRY(dalloc(b, &b->s.evals32, b->W));
    HIPCHK(hipMemset(b->s.evals32, 0, b->W * sizeof(unsigned long long)));
    b->s.act = act;
    b->s.nact = nact;
    b->hfrozen.assign(b->A, 0);
    b->hact.assign(b->A, 0);
    b->hnact.assign(b->W, 0);
    TRY(dalloc(b, &b->s.ws, b->W));
    HIPCHK(hipMemset(b->s.ws, 0, b->W * sizeof(WalkerState)));
    TRY(dalloc(b, &b->zflag, b->W));
    TRY(dalloc(b, &b->zmark, b->A));
    TRY(dalloc(b, &b->zlist, b->A));
    TRY(dalloc(b, &b->nz, b->W));
    // a list chunk holds at least 32 slots (256 threads at 8 lanes per atom)
    for (int k = 0; k < 4; ++k) TRY(dalloc(b, &b->wl_tab[k], (size_t)b->A / 32 + b->W + 1));
    TRY(dalloc(b, &b->wl_count, 4));
    HIPCHK(hipMemset(b->wl_count, 0, 4 * sizeof(int)));
    {
        // persistent grids: every CU, as many staged workgroups as its 160 KB
        // of LDS holds (at most 2), 8 plain ones; multiples of the 8 XCDs
        int dev = 0, ncu = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const size_t lds = (size_t)b->d.stage_bytes + sizeof(SCPairDev) * AKMC_MAXPAIR;
        const int res = std::max(1, std::min(2, (int)(163840 / std::max<size_t>(lds, 1))));
        b->wl_grid_staged = ((ncu * res + 7) / 8) * 8;
        b->wl_grid_plain = ((ncu * 8 + 7) / 8) * 8;
    }
    HIPCHK(hipMemset(b->zflag, 0, b->W * sizeof(int)));
    HIPCHK(hipMemset(b->nz, 0, b->W * sizeof(int)));
    b->hzflag.assign(b->W, 0);
    // X starts at the resident positions
    size_t A = b->A;
    HIPCHK(hipMemcpy(b->s.X, b->d.x, A * sizeof(double), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(b->s.X + A, b->d.y, A * sizeof(double), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(b->s.X + 2 * A, b->d.z, A * sizeof(double), hipMemcpyDeviceToDevice));
    b->hws.assign(b->W, WalkerState{});
    TRY(upload_active(b, nullptr));
    HIPCHK(hipHostMalloc((void **)&b->hws_pin, b->W * sizeof(WalkerState), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&b->ov_pin, sizeof(int), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&b->st_pin, b->W * sizeof(StartItem), hipHostMallocDefault));
    TRY(dalloc(b, &b->st_dev, b->W));
    HIPCHK(hipEventCreateWithFlags(&b->stage_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->st_ev, hipEventDisableTiming));
    b->search_alloc = true;
    return AKMC_SUCCESS;
}

void drop_graph(akmc_batch_t b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->graph_steps = 0;
}

// active zones apply to minimum-image Sutton-Chen fields
bool zone_usable(akmc_batch_t b)
{
    const FieldDev &h = b->f->host;
    return b->zone_on && b->zflag && h.kind == 0 && !h.eam && h.min_image;
}

// the batch view of the search micro-steps: the plain one plus the pair
// energies in pass 1, the search variants and the zones
BatchDev step_dev(akmc_batch_t b)
{
    BatchDev d = b->d;
    const FieldDev &h = b->f->host;
    if (h.kind == 0 && !h.eam && h.min_image) {
        d.pe1 = 1;
        for (int k = 0; k < 4; ++k) d.wl_tab[k] = b->wl_tab[k];
        d.wl_count = b->wl_count;
        d.wl_grid_staged = b->wl_grid_staged;
        d.wl_grid_plain = b->wl_grid_plain;
        if (b->sv[0]) {
            d.lpa = b->sv[0];
            d.unroll = b->sv[1];
            d.staged = b->sv[2];
        }
        if (b->sv[3]) {
            d.lpa2 = b->sv[3];
            d.unroll2 = b->sv[4];
            d.staged2 = b->sv[5];
        }
        if (b->sv[6]) {
            d.lpa64 = b->sv[6];
            d.unroll64 = b->sv[7];
            d.staged64 = b->sv[8];
        }
    }
    if (zone_usable(b)) {
        d.zfrozen = b->s.frozen;
        d.zflag = b->zflag;
        d.zmark = b->zmark;
        d.zlist = b->zlist;
        d.nz = b->nz;
        d.zact = b->s.act;
        d.znact = b->s.nact;
    }
    return d;
}

// one micro-step; step j of a run chunk records the kernel-timer events
void enqueue_micro_step(akmc_batch_t b, int j)
{
    // FP64 pair terms run when a walker may ask for them: verification or
    // the closing FP64 energy of a relax
    int prec = batch_mode(b) | ((b->s.verify_nl > 0 || b->s.final64) ? 2 : 0);
    const BatchDev d = step_dev(b);
    if (b->ktimer && (prec & 1) && b->f->host.kind == 0) {
        for (int ph = 0; ph < 2; ++ph) akmc_launch_phase(d, b->f->host, prec, 0, ph, b->stream);
        (void)hipEventRecord(b->kev[0][j], b->stream);
        akmc_launch_phase(d, b->f->host, 1, 0, 2, b->stream);
        (void)hipEventRecord(b->kev[1][j], b->stream);
        if (prec & 2) akmc_launch_phase(d, b->f->host, 2 | 4, 0, 2, b->stream);  // FP64 part of a mixed step
        akmc_launch_phase(d, b->f->host, prec, 0, 3, b->stream);
    } else {
        akmc_launch_forces_impl(d, b->f->host, prec, 0, b->stream);
    }
    // walkers of tens of thousands of atoms: 1024-thread state machines
    if (b->logic_wide) akmc_launch_search_logic_wide(d, b->s, b->stream);
    else akmc_launch_search_logic(d, b->s, b->stream);
}

int fetch_states(akmc_batch_t b)
{
    HIPCHK(hipMemcpyAsync(b->hws.data(), b->s.ws, b->W * sizeof(WalkerState), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}

// mark walkers that are done (or idle) inactive so the force kernels skip them
int refresh_active(akmc_batch_t b, int *n_running)
{
    std::vector<int> act(b->W);
    int run = 0;
    for (int w = 0; w < b->W; ++w) {
        const WalkerState &ws = b->hws[w];
        const bool fp64 = ws.kind == WK_VERIFY || ws.state == ST_R_E64;
        act[w] = (ws.kind != WK_IDLE && ws.state != ST_DONE) ? (fp64 ? 2 : batch_mode(b)) : 0;
        run += act[w];
    }
    HIPCHK(hipMemcpyAsync(b->d.active, act.data(), b->W * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    *n_running = run;
    return AKMC_SUCCESS;
}

// the host may rewrite a pinned staging buffer once its last transfer,
// recorded on ev, is done
int wait_staging(hipEvent_t ev, bool &pending)
{
    if (pending) HIPCHK(hipEventSynchronize(ev));
    pending = false;
    return AKMC_SUCCESS;
}

// The listed walkers' fresh states go to the device in one pinned transfer
// and one scatter kernel, with, per walker: a fresh zone when it has frozen
// atoms (membership reset, and a full evaluation first in each precision),
// a list rebuild on its first micro-step (zones on or off alike, so both
// list the same neighbours in the same order) and its active flag.  Only the
// started walkers' device state is written, so a start may be queued behind
// a run chunk in flight (it then returns at once); otherwise it waits.
int start_walkers(akmc_batch_t b, const int *walkers, int kind, double dt0, double alpha0)
{
    TRY(wait_staging(b->st_ev, b->st_pending));
    int k = 0;
    for (int w = 0; w < b->W; ++w) {
        if (walkers && !walkers[w]) continue;
        WalkerState ws{};
        ws.kind = kind;
        ws.state = kind == WK_RELAX ? ST_R_INIT : (kind == WK_ART ? ST_A_L0 : ST_D_F0);
        ws.it = (kind == WK_DIMER) ? 1 : 0;
        ws.status = AKMC_FAILED;
        ws.dt = dt0;
        ws.alpha = alpha0;
        // Relax.cpp:49 and ART.cpp:426 start nStep at 1
        ws.nstep = (kind == WK_RELAX || kind == WK_ART) ? 1 : 0;
        ws.inloop = 0;
        ws.verified_state = 0;
        b->hws[w] = ws;
        const int m = b->hnact[w];
        StartItem &it = b->st_pin[k++];
        it.walker = w;
        it.zflag = (m > 0 && m < b->hw[w].n) ? (AKMC_ZONE_ON | AKMC_ZONE_FULL32 | AKMC_ZONE_FULL64 | AKMC_ZONE_RESET)
                                             : 0;
        it.active = batch_mode(b);
        it.pad = 0;
        it.ws = ws;
        b->hzflag[w] = it.zflag;
        if (b->inflight) b->started_inflight.emplace_back(w, ws);
    }
    if (k == 0) return AKMC_SUCCESS;
    HIPCHK(hipMemcpyAsync(b->st_dev, b->st_pin, k * sizeof(StartItem), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipEventRecord(b->st_ev, b->stream));
    b->st_pending = true;
    akmc_launch_start_walkers(b->d, b->s, b->st_dev, k, b->zflag, b->hv_on ? b->harvested : nullptr, b->stream);
    akmc_launch_probe_init(b->d, b->s, b->stream);
    HIPCHK(hipGetLastError());
    if (!b->inflight) HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}

int ensure_lanczos(akmc_batch_t b)
{
    if (b->s.Q) return AKMC_SUCCESS;
    TRY(dalloc(b, &b->s.Q, (size_t)AKMC_LZ_MAX * 3 * b->A));
    TRY(dalloc(b, &b->s.lz, (size_t)b->W * 3 * AKMC_LZ_MAX));
    return AKMC_SUCCESS;
}

// grow a pinned host / device staging pair to n bytes
int ensure_stage(akmc_batch_t b, size_t n)
{
    if (n <= b->stage_n) return AKMC_SUCCESS;
    if (b->stage_pin) HIPCHK(hipHostFree(b->stage_pin));
    if (b->stage_dev) HIPCHK(hipFree(b->stage_dev));
    b->stage_pin = nullptr;
    b->stage_dev = nullptr;
    b->stage_n = 0;
    HIPCHK(hipHostMalloc((void **)&b->stage_pin, n, hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&b->stage_dev, n));
    b->stage_n = n;
    return AKMC_SUCCESS;
}

// per-walker start data of a saddle search: positions (optional), a
// normalised mode with frozen components removed, zero velocity, the frozen
// mask and the free-atom list.  The selected walkers' data are packed into
// one pinned buffer, copied in one transfer and scattered by one kernel.
int load_walkers(akmc_batch_t b, const int *walkers, const double *x0, const double *mode, const int *frozen)
{
    const size_t A = b->A;
    TRY(wait_staging(b->stage_ev, b->stage_pending));
    std::vector<LoadItem> items;
    size_t dwords = 0, iwords = 0;
    for (int w = 0; w < b->W; ++w) {
        if (walkers && !walkers[w]) continue;
        const int off = b->hw[w].off, n = b->hw[w].n;
        if (frozen) std::copy(frozen + off, frozen + off + n, b->hfrozen.begin() + off);
        LoadItem it{};
        it.walker = w;
        it.dbl = (long long)dwords;
        it.ints = (long long)iwords;
        dwords += (x0 ? 6 : 3) * (size_t)n;
        iwords += 2 * (size_t)n;
        items.push_back(it);
    }
    if (items.empty()) return AKMC_SUCCESS;
    const size_t ibytes = items.size() * sizeof(LoadItem);
    const size_t dbytes = dwords * sizeof(double), nbytes = iwords * sizeof(int);
    TRY(ensure_stage(b, ibytes + dbytes + nbytes + 64));
    char *base = (char *)b->stage_pin;
    double *D = (double *)(base + ibytes);
    int *I = (int *)(base + ibytes + dbytes);
    for (LoadItem &it : items) {
        const int w = it.walker, off = b->hw[w].off, n = b->hw[w].n;
        const int *fr = b->hfrozen.data() + off;
        double *nw = D + it.dbl;
        double nn = 0.0;
        for (int c = 0; c < 3; ++c)
            for (int i = 0; i < n; ++i) {
                double v = fr[i] ? 0.0 : mode[c * A + off + i];
                nw[(size_t)c * n + i] = v;
                nn += v * v;
            }
        if (nn <= 0.0) return fail(AKMC_EINVAL, "walker %d: zero start mode", w);
        nn = 1.0 / std::sqrt(nn);
        for (size_t k = 0; k < 3 * (size_t)n; ++k) nw[k] *= nn;
        it.has_x = x0 ? 1 : 0;
        if (x0)
            for (int c = 0; c < 3; ++c) std::memcpy(nw + (size_t)(3 + c) * n, x0 + c * A + off, n * sizeof(double));
        // frozen mask and free-atom list (global ids)
        int *iw = I + it.ints;
        int m = 0;
        for (int i = 0; i < n; ++i) {
            iw[i] = fr[i];
            if (!fr[i]) {
                iw[n + m++] = off + i;
                b->hact[off + m - 1] = off + i;
            }
        }
        it.nact = m;
        b->hnact[w] = m;
    }
    std::memcpy(base, items.data(), ibytes);
    HIPCHK(hipMemcpyAsync(b->stage_dev, b->stage_pin, ibytes + dbytes + nbytes, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipEventRecord(b->stage_ev, b->stream));
    b->stage_pending = true;
    char *dbase = (char *)b->stage_dev;
    akmc_launch_load_walkers(b->d, b->s, (const LoadItem *)dbase, (int)items.size(), (const double *)(dbase + ibytes),
                             (const int *)(dbase + ibytes + dbytes), b->stream);
    HIPCHK(hipGetLastError());
    // queued behind a chunk in flight, the load returns at once
    if (!b->inflight) HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}

int copy_walkers_out(akmc_batch_t b, const double *src, int nw, const int *walkers, double *out);
int copy_walkers_out3(akmc_batch_t b, const double *const src[3], int nw, const int *walkers, double *out);

int capture_graph(akmc_batch_t b, int steps)
{
    drop_graph(b);
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < steps; ++k) enqueue_micro_step(b, k);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess) return fail(AKMC_EHIP, "graph capture failed: %s", hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->graph_steps = steps;
    b->graph_args = b->s;
    b->graph_d = step_dev(b);
    b->graph_prec = batch_mode(b);
    return AKMC_SUCCESS;
}

}  // namespace

// ======================================================================
extern "C" {

const char *akmc_last_error(void) { return g_err.c_str(); }

int akmc_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int akmc_version(void) { return 1; }

int akmc_field_create(const char *text, akmc_field_t *out)
{
    if (!out) return fail(AKMC_EINVAL, "null output handle");
    akmc_field_t f = new akmc_field_s();
    int rc = parse_potentials(f, text);
    if (rc != AKMC_SUCCESS) {
        delete f;
        return rc;
    }
    build_tables(f);
    *out = f;
    return AKMC_SUCCESS;
}

void akmc_field_destroy(akmc_field_t f)
{
    if (!f) return;
    if (f->single) batch_free(f->single);
    if (f->dev) (void)hipFree(f->dev);
    if (f->eam_dev) (void)hipFree(f->eam_dev);
    if (f->dsf_dev) (void)hipFree(f->dsf_dev);
    delete f;
}

int akmc_field_kind(akmc_field_t f) { return f ? f->kind : AKMC_EINVAL; }

int akmc_field_set_ewald(akmc_field_t f, double eta, int gx, int gy, int gz)
{
    if (!f || eta <= 0 || gx < 1 || gy < 1 || gz < 1) return fail(AKMC_EINVAL, "bad Ewald parameters");
    f->explicit_ewald = true;
    f->eta = eta;
    f->gcell[0] = gx;
    f->gcell[1] = gy;
    f->gcell[2] = gz;
    return AKMC_SUCCESS;
}

int akmc_field_set_skin(akmc_field_t f, double skin)
{
    if (!f || skin < 0) return fail(AKMC_EINVAL, "bad skin");
    f->skin = skin;
    build_tables(f);
    return AKMC_SUCCESS;
}

int akmc_field_set_list_density(akmc_field_t f, double atoms_per_A3)
{
    if (!f || !(atoms_per_A3 >= 0.0) || atoms_per_A3 > 10.0) return fail(AKMC_EINVAL, "bad list density");
    f->list_density = atoms_per_A3;
    return AKMC_SUCCESS;
}

int akmc_field_set_precision(akmc_field_t f, int bits)
{
    if (!f || (bits != 32 && bits != 64)) return fail(AKMC_EINVAL, "precision must be 32 or 64");
    f->fp64 = bits == 64;
    return AKMC_SUCCESS;
}

int akmc_field_setup(akmc_field_t f, const double lat[9], int max_atoms)
{
    if (!f || !lat || max_atoms <= 0) return fail(AKMC_EINVAL, "bad setup arguments");
    TRY(check_cell(f, lat));
    return ensure_single(f, max_atoms, lat);
}

int akmc_field_num_kvectors(akmc_field_t f) { return f ? f->nk_last : 0; }

}  // extern "C"

int akmc_field_view(akmc_field_s *f, FieldView *v)
{
    if (!f || !v) return fail(AKMC_EINVAL, "bad arguments");
    v->dev = f->dev;
    v->host = &f->host;
    v->rc = f->rc;
    v->skin = f->skin;
    v->kind = f->kind;
    v->min_image = f->min_image ? 1 : 0;
    v->eam = f->eam.empty() ? 0 : 1;
    v->fp64 = f->fp64;
    return AKMC_SUCCESS;
}

extern "C" {

int akmc_field_forces(akmc_field_t f, double *x, double *y, double *z, double *fx, double *fy, double *fz,
                      double stress[6], const double lat[9], const double *charge, const int *type,
                      const int *frozen, int n, akmc_energy *e)
{
    (void)frozen;  // unused by the reference's force call as well
    if (!f || !x || !y || !z || !fx || !fy || !fz || !type || n <= 0) return fail(AKMC_EINVAL, "bad arguments");
    if (f->kind == 1 && f->do_ewald && !charge) return fail(AKMC_EINVAL, "ionic field needs charges");
    TRY(ensure_single(f, n, lat));
    akmc_batch_t b = f->single;
    TRY(upload(b, x, y, z, type, charge));
    TRY(run_forces(b, stress != nullptr));
    TRY(check_overflow(b));
    HIPCHK(hipMemcpyAsync(fx, b->d.fx, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(fy, b->d.fy, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(fz, b->d.fz, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    TRY(fetch_out(b));
    if (e) to_energy(b->hout[0], e);
    if (stress)
        for (int c = 0; c < 6; ++c) stress[c] = b->hout[0].stress[c];
    host_wrap(f->min_image, lat, n, x, y, z);
    return AKMC_SUCCESS;
}

int akmc_field_energy(akmc_field_t f, double *x, double *y, double *z, const double lat[9], const double *charge,
                      const int *type, int n, akmc_energy *e)
{
    if (!f || !x || !y || !z || !type || n <= 0) return fail(AKMC_EINVAL, "bad arguments");
    if (f->kind == 1 && f->do_ewald && !charge) return fail(AKMC_EINVAL, "ionic field needs charges");
    TRY(ensure_single(f, n, lat));
    akmc_batch_t b = f->single;
    TRY(upload(b, x, y, z, type, charge));
    // this launch's own copy of the batch descriptor carries the energy-only
    // flag: no force pass, and the energy-only short-range forms of the
    // two-body table (TwoBody.h:81-150: Morse adds C12/r^12).  Tabulated EAM
    // keeps its force pass (its pair energy comes from it).
    BatchDev d = b->d;
    d.energy_only = 1;
    akmc_launch_forces_impl(d, f->host, batch_mode(b), 0, b->stream);
    HIPCHK(hipGetLastError());
    TRY(check_overflow(b));
    TRY(fetch_out(b));
    if (e) to_energy(b->hout[0], e);
    host_wrap(f->min_image, lat, n, x, y, z);
    // the energy path of a tabulated EAM field doubles the pair term
    // (calculateManyBodyPairEnergy, potentials/metal/ManyBody.cpp:444-450)
    if (e && f->kind == 0 && !f->eam.empty()) {
        e->vdw *= 2.0;
        e->total = e->many + e->vdw;
    }
    return AKMC_SUCCESS;
}

int akmc_batch_create(akmc_field_t f, int nwalkers, const int *natoms, const double *lat, akmc_batch_t *out)
{
    if (!f || nwalkers <= 0 || !natoms || !lat || !out) return fail(AKMC_EINVAL, "bad batch arguments");
    akmc_batch_t b = new akmc_batch_s();
    int rc = batch_build(f, nwalkers, natoms, lat, b);
    if (rc != AKMC_SUCCESS) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return AKMC_SUCCESS;
}

void akmc_batch_destroy(akmc_batch_t b) { batch_free(b); }

int akmc_batch_upload(akmc_batch_t b, const double *x, const double *y, const double *z, const int *type,
                      const double *charge, const int *frozen)
{
    (void)frozen;
    if (!b || !x || !y || !z || !type) return fail(AKMC_EINVAL, "bad arguments");
    return upload(b, x, y, z, type, charge);
}

int akmc_batch_download(akmc_batch_t b, double *x, double *y, double *z, double *fx, double *fy, double *fz)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    size_t A = b->A;
    if (x) HIPCHK(hipMemcpyAsync(x, b->d.x, A * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (y) HIPCHK(hipMemcpyAsync(y, b->d.y, A * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (z) HIPCHK(hipMemcpyAsync(z, b->d.z, A * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (fx) HIPCHK(hipMemcpyAsync(fx, b->d.fx, A * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (fy) HIPCHK(hipMemcpyAsync(fy, b->d.fy, A * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (fz) HIPCHK(hipMemcpyAsync(fz, b->d.fz, A * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}

int akmc_batch_forces(akmc_batch_t b, akmc_energy *energies)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    TRY(run_forces(b, 0));
    TRY(check_overflow(b));
    TRY(fetch_out(b));
    if (energies)
        for (int w = 0; w < b->W; ++w) to_energy(b->hout[w], energies + w);
    return AKMC_SUCCESS;
}

int akmc_batch_forces_repeat(akmc_batch_t b, int nrep, double *ms)
{
    if (!b || nrep <= 0) return fail(AKMC_EINVAL, "bad arguments");
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    for (int r = 0; r < nrep; ++r) TRY(run_forces(b, 0));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, b->ev0, b->ev1));
    if (ms) *ms = t;
    return check_overflow(b);
}

int akmc_batch_set_variant(akmc_batch_t b, int lanes, int unroll)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8)
        return fail(AKMC_EINVAL, "lanes per atom must be 1, 2, 4 or 8, got %d", lanes);
    if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8)
        return fail(AKMC_EINVAL, "unroll must be 1, 2, 4 or 8, got %d", unroll);
    if (b->d.lpa != lanes || b->d.unroll != unroll) drop_graph(b);  // the captured graph holds the old kernels
    b->d.lpa = lanes;
    b->d.unroll = unroll;
    return AKMC_SUCCESS;
}

int akmc_batch_set_staging(akmc_batch_t b, int on)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    if (on && !b->stage_ok)
        return fail(AKMC_EUNSUPPORTED,
                    "LDS staging needs a many-body field whose largest walker's gather records "
                    "(%d bytes) fit one workgroup's LDS", b->d.stage_bytes);
    if (b->d.staged != (on ? 1 : 0)) drop_graph(b);
    b->d.staged = on ? 1 : 0;
    return AKMC_SUCCESS;
}

int akmc_batch_staging(akmc_batch_t b, int *on, int *possible)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    if (on) *on = b->d.staged;
    if (possible) *possible = b->stage_ok ? 1 : 0;
    return AKMC_SUCCESS;
}

int akmc_batch_autotune(akmc_batch_t b, int nrep, int best[3], double ms[AKMC_NVARIANTS])
{
    if (!b || nrep <= 0) return fail(AKMC_EINVAL, "bad arguments");
    static const int lanes[4] = {1, 2, 4, 8}, unroll[4] = {1, 2, 4, 8};
    double t[AKMC_NVARIANTS];
    TRY(run_forces(b, 0));  // lists current before timing
    const int nv = b->stage_ok ? AKMC_NVARIANTS : AKMC_NVARIANTS / 2;
    for (int v = 0; v < AKMC_NVARIANTS; ++v) {
        t[v] = -1.0;
        if (v >= nv) continue;
        TRY(akmc_batch_set_staging(b, v / 16));
        TRY(akmc_batch_set_variant(b, lanes[v / 4 % 4], unroll[v % 4]));
        TRY(run_forces(b, 0));  // first launch of this instantiation
        TRY(akmc_batch_forces_repeat(b, nrep, &t[v]));
    }
    if (ms)
        for (int v = 0; v < AKMC_NVARIANTS; ++v) ms[v] = t[v];
    int k = 0;
    for (int v = 1; v < nv; ++v)
        if (t[v] < t[k]) k = v;
    TRY(akmc_batch_set_staging(b, k / 16));
    TRY(akmc_batch_set_variant(b, lanes[k / 4 % 4], unroll[k % 4]));
    if (best) {
        best[0] = lanes[k / 4 % 4];
        best[1] = unroll[k % 4];
        best[2] = k / 16;
    }
    return AKMC_SUCCESS;
}

int akmc_batch_total_atoms(akmc_batch_t b) { return b ? b->A : 0; }

long long akmc_batch_rebuilds(akmc_batch_t b)
{
    if (!b) return 0;
    unsigned long long v = 0;
    if (hipMemcpy(&v, akmc_rebuild_counter_ptr(), sizeof v, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return (long long)(v - b->rebuild_base);
}


void akmc_relax_defaults(akmc_relax_params *p)
{
    if (!p) return;
    p->maxIter = 3000;
    p->normTol = 1.0e-5;
    p->forceTol = 1.0e-4;
    p->N_min = 5;
    p->initTimeStep = 0.0005;
    p->alphaDecrease = 0.99;
    p->stepIncrease = 1.1;
    p->stepDecrease = 0.5;
    p->alphaStart = 0.1;
    p->maxTimeStep = 0.00075;
    p->timeStepUnit = 9648.530821;
    p->method = AKMC_RELAX_FIRE;
    p->verletIntegration = 0;
    p->N_delaySteps = 20;
    p->N_maxUphill = 50;
    p->initialDelay = 1;
    p->minTimeStep = 0.0001;
}

void akmc_dimer_defaults(akmc_dimer_params *p)
{
    if (!p) return;
    p->deltaTheta = 0.005;
    p->dimerDisplacement = 0.01;
    p->maxDeltaX = 0.5;
    p->maxRotations = 20;
    p->maxParIter = 1000;
    p->maxEigenvalue = 0.0;
    p->normTol = 0.1;
    p->forceTol = 1.0e-2;
    p->N_min = 5;
    p->alphaDecrease = 0.99;
    p->stepIncrease = 1.1;
    p->stepDecrease = 0.5;
    p->alphaStart = 0.1;
    p->initTimeStep = 0.005;
    p->maxTimeStep = 0.01;
    p->maxStep = 0.1;
    p->timeStepUnit = 9648.530821;
}

void akmc_art_defaults(akmc_art_params *p)
{
    if (!p) return;
    p->numLanczosVectors = 20;
    p->deltaX = 0.01;
    p->maxEigenvalue = 0.0;
    p->eigValTol = 0.001;
    p->parDamp = 0.5;
    p->maxParIter = 1000;
    p->normTol = 1.0e-4;
    p->forceTol = 1.0e-3;
    p->N_min = 5;
    p->initTimeStep = 0.005;
    p->alphaDecrease = 0.99;
    p->stepIncrease = 1.1;
    p->stepDecrease = 0.5;
    p->alphaStart = 0.1;
    p->maxTimeStep = 0.01;
    p->timeStepUnit = 9648.530821;
}

int akmc_batch_set_frozen(akmc_batch_t b, const int *frozen)
{
    if (!b || !frozen) return fail(AKMC_EINVAL, "bad arguments");
    TRY(search_alloc(b));
    HIPCHK(hipMemcpy((void *)b->s.frozen, frozen, b->A * sizeof(int), hipMemcpyHostToDevice));
    std::copy(frozen, frozen + b->A, b->hfrozen.begin());
    return upload_active(b, nullptr);
}

int akmc_batch_start_relax(akmc_batch_t b, const int *walkers, const akmc_relax_params *p)
{
    if (!b || !p) return fail(AKMC_EINVAL, "bad arguments");
    TRY(search_alloc(b));
    if (p->method != AKMC_RELAX_FIRE && p->method != AKMC_RELAX_FIRE2 && p->method != AKMC_RELAX_NONE)
        return fail(AKMC_EINVAL, "relax method %d is not one of AKMC_RELAX_FIRE, _FIRE2, _NONE", p->method);
    RelaxParamsDev &r = b->s.rp;
    r.normTol = p->normTol;
    r.forceTol = p->forceTol;
    r.initTimeStep = p->initTimeStep;
    r.maxTimeStep = p->maxTimeStep;
    r.alphaStart = p->alphaStart;
    r.alphaDecrease = p->alphaDecrease;
    r.stepIncrease = p->stepIncrease;
    r.stepDecrease = p->stepDecrease;
    r.timeStepUnit = p->timeStepUnit;
    r.minTimeStep = p->minTimeStep;
    r.maxIter = p->maxIter;
    r.N_min = p->N_min;
    r.method = p->method;
    r.verlet = p->verletIntegration ? 1 : 0;
    r.N_delaySteps = p->N_delaySteps;
    r.N_maxUphill = p->N_maxUphill;
    r.initialDelay = p->initialDelay ? 1 : 0;
    // velocities start at zero
    if (!walkers) {
        HIPCHK(hipMemset(b->s.V, 0, 3 * (size_t)b->A * sizeof(double)));
    } else {
        for (int w = 0; w < b->W; ++w)
            if (walkers[w])
                for (int c = 0; c < 3; ++c)
                    HIPCHK(hipMemset(b->s.V + (size_t)c * b->A + b->hw[w].off, 0, b->hw[w].n * sizeof(double)));
    }
    return start_walkers(b, walkers, WK_RELAX, p->initTimeStep, p->alphaStart);
}

int akmc_batch_start_dimer(akmc_batch_t b, const int *walkers, const double *x0, const double *mode,
                           const int *frozen, const akmc_dimer_params *p)
{
    if (!b || !p || !mode) return fail(AKMC_EINVAL, "bad arguments");
    TRY(search_alloc(b));
    DimerParamsDev &d = b->s.dp;
    d.deltaTheta = p->deltaTheta;
    d.dimerDisplacement = p->dimerDisplacement;
    d.maxEigenvalue = p->maxEigenvalue;
    d.normTol = p->normTol;
    d.forceTol = p->forceTol;
    d.alphaStart = p->alphaStart;
    d.alphaDecrease = p->alphaDecrease;
    d.stepIncrease = p->stepIncrease;
    d.stepDecrease = p->stepDecrease;
    d.initTimeStep = p->initTimeStep;
    d.maxTimeStep = p->maxTimeStep;
    d.maxStep = p->maxStep;
    d.timeStepUnit = p->timeStepUnit;
    d.maxRotations = p->maxRotations;
    d.maxParIter = p->maxParIter;
    d.N_min = p->N_min;
    TRY(load_walkers(b, walkers, x0, mode, frozen));
    return start_walkers(b, walkers, WK_DIMER, p->initTimeStep, p->alphaStart);
}

int akmc_batch_start_art(akmc_batch_t b, const int *walkers, const double *x0, const double *mode,
                         const int *frozen, const akmc_art_params *p)
{
    if (!b || !p || !mode) return fail(AKMC_EINVAL, "bad arguments");
    if (p->numLanczosVectors < 2 || p->numLanczosVectors > AKMC_LZ_MAX)
        return fail(AKMC_EINVAL, "numLanczosVectors must be in [2, %d]", AKMC_LZ_MAX);
    TRY(search_alloc(b));
    TRY(ensure_lanczos(b));
    ARTParamsDev &a = b->s.ap;
    a.deltaX = p->deltaX;
    a.maxEigenvalue = p->maxEigenvalue;
    a.eigValTol = p->eigValTol;
    a.parDamp = p->parDamp;
    a.normTol = p->normTol;
    a.forceTol = p->forceTol;
    a.initTimeStep = p->initTimeStep;
    a.alphaDecrease = p->alphaDecrease;
    a.stepIncrease = p->stepIncrease;
    a.stepDecrease = p->stepDecrease;
    a.alphaStart = p->alphaStart;
    a.maxTimeStep = p->maxTimeStep;
    a.timeStepUnit = p->timeStepUnit;
    a.numLanczosVectors = p->numLanczosVectors;
    a.maxParIter = p->maxParIter;
    a.N_min = p->N_min;
    TRY(load_walkers(b, walkers, x0, mode, frozen));
    return start_walkers(b, walkers, WK_ART, p->initTimeStep, p->alphaStart);
}

}  // extern "C"

namespace {

// launches one chunk of `steps` micro-steps and the copies of the walker
// states and the overflow flag that its wait reads
int launch_chunk(akmc_batch_t b, int steps, int running)
{
    b->inflight_steps = 0;
    b->inflight = 1;
    if (running == 0) return AKMC_SUCCESS;
    // k micro-steps as one graph launch; the state kernels keep the
    // per-walker active flags current (0 when a walker ends, FP64 while
    // it verifies), so the host only reads states and the overflow flag
    if (b->ktimer && steps > akmc_batch_s::kTimerSteps) {
        b->inflight = 0;
        return fail(AKMC_EINVAL, "sync_every is limited to %d while the kernel timer runs", akmc_batch_s::kTimerSteps);
    }
    // HIP events recorded inside a replayed graph cannot be timed, so the
    // kernel timer launches the micro-steps one by one
    if (b->use_graphs && !b->ktimer) {
        const BatchDev bd = step_dev(b);
        if (!b->gexec || b->graph_steps != steps || b->graph_prec != batch_mode(b) ||
            std::memcmp(&b->graph_args, &b->s, sizeof(SearchDev)) != 0 ||
            std::memcmp(&b->graph_d, &bd, sizeof(BatchDev)) != 0)
            TRY(capture_graph(b, steps));
        HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    } else {
        for (int j = 0; j < steps; ++j) enqueue_micro_step(b, j);
        HIPCHK(hipGetLastError());
    }
    if (b->hv_on) {
        // the buffer the host read two waits ago
        b->hv_cur ^= 1;
        akmc_launch_harvest_pack(b->d, b->s, b->harvested, b->hv_dev[b->hv_cur], b->stream);
        HIPCHK(hipGetLastError());
        b->hv_inflight = b->hv_cur;
    }
    HIPCHK(hipMemcpyAsync(b->hws_pin, b->s.ws, b->W * sizeof(WalkerState), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(b->ov_pin, b->d.overflow, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    b->inflight_steps = steps;
    return AKMC_SUCCESS;
}

}  // namespace

extern "C" {

int akmc_batch_run_async(akmc_batch_t b, int steps)
{
    if (!b || !b->search_alloc || steps < 1) return fail(AKMC_EINVAL, "bad arguments");
    if (b->inflight) return fail(AKMC_EINVAL, "a run chunk is already in flight");
    b->s.default_mode = batch_mode(b);  // the mode the state kernels give walkers they move on
    TRY(fetch_states(b));
    int running = 0;
    TRY(refresh_active(b, &running));
    return launch_chunk(b, steps, running);
}

int akmc_batch_run_next(akmc_batch_t b, int steps)
{
    if (!b || !b->search_alloc || steps < 1) return fail(AKMC_EINVAL, "bad arguments");
    if (b->inflight) return fail(AKMC_EINVAL, "a run chunk is already in flight");
    b->s.default_mode = batch_mode(b);
    // the host's copy of the states is the last wait's, with the starts
    // made since; the device keeps the active flags
    int running = 0;
    for (int w = 0; w < b->W; ++w) running += b->hws[w].kind != WK_IDLE && b->hws[w].state != ST_DONE;
    return launch_chunk(b, steps, running);
}

int akmc_batch_run_wait(akmc_batch_t b, int *steps_done, int *running)
{
    if (!b || !b->inflight) return fail(AKMC_EINVAL, "no run chunk in flight");
    b->inflight = 0;
    const int k = b->inflight_steps;
    int run = 0;
    if (k > 0) {
        HIPCHK(hipStreamSynchronize(b->stream));
        if (*b->ov_pin) return overflow_error(b, *b->ov_pin);
        if (b->ktimer && batch_mode(b) == 1 && b->f->host.kind == 0)
            for (int j = 0; j < k; ++j) {
                float t = 0.0f;
                HIPCHK(hipEventElapsedTime(&t, b->kev[0][j], b->kev[1][j]));
                b->kms += t;
                b->klaunches += 1;
            }
        std::memcpy(b->hws.data(), b->hws_pin, b->W * sizeof(WalkerState));
        if (b->hv_on) b->hv_last = b->hv_inflight;
    } else {
        HIPCHK(hipStreamSynchronize(b->stream));  // starts queued since the launch
    }
    // walkers started behind the chunk have not run yet
    for (const auto &p : b->started_inflight) b->hws[p.first] = p.second;
    b->started_inflight.clear();
    for (int w = 0; w < b->W; ++w) run += b->hws[w].kind != WK_IDLE && b->hws[w].state != ST_DONE;
    if (steps_done) *steps_done = k;
    if (running) *running = run;
    return AKMC_SUCCESS;
}

int akmc_batch_set_harvest(akmc_batch_t b, int on)
{
    if (!b || !b->search_alloc) return fail(AKMC_EINVAL, "no search has been started");
    if (b->inflight) return fail(AKMC_EINVAL, "a run chunk is in flight");
    if (on && !b->harvested) {
        TRY(dalloc(b, &b->harvested, b->W));
        for (int k = 0; k < 2; ++k) {
            HIPCHK(hipHostMalloc((void **)&b->hv_pin[k], 3 * (size_t)b->A * sizeof(double), hipHostMallocDefault));
            HIPCHK(hipHostGetDevicePointer((void **)&b->hv_dev[k], b->hv_pin[k], 0));
        }
    }
    // walkers that have already ended are harvested by the next chunk
    if (on) HIPCHK(hipMemsetAsync(b->harvested, 0, b->W * sizeof(int), b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->hv_on = on ? 1 : 0;
    b->hv_last = -1;
    return AKMC_SUCCESS;
}

int akmc_batch_get_harvested_positions(akmc_batch_t b, int nw, const int *walkers, double *out)
{
    if (!b || nw < 0 || (nw && (!walkers || !out))) return fail(AKMC_EINVAL, "bad arguments");
    if (!b->hv_on || b->hv_last < 0) return fail(AKMC_EINVAL, "no harvested chunk has been waited for");
    const double *src = b->hv_pin[b->hv_last];
    const size_t A = b->A;
    size_t o = 0;
    for (int k = 0; k < nw; ++k) {
        const int w = walkers[k];
        if (w < 0 || w >= b->W) return fail(AKMC_EINVAL, "walker %d out of range", w);
        const int off = b->hw[w].off, n = b->hw[w].n;
        for (int c = 0; c < 3; ++c, o += n) std::memcpy(out + o, src + c * A + off, n * sizeof(double));
    }
    return AKMC_SUCCESS;
}

int akmc_batch_run(akmc_batch_t b, int max_steps, int sync_every, int *steps_done)
{
    if (!b || !b->search_alloc || max_steps < 0 || sync_every < 1) return fail(AKMC_EINVAL, "bad arguments");
    int done = 0, running = 1;
    while (running > 0 && done < max_steps) {
        int k = std::min(sync_every, max_steps - done), got = 0;
        TRY(akmc_batch_run_async(b, k));
        TRY(akmc_batch_run_wait(b, &got, &running));
        if (got == 0) break;
        done += got;
    }
    if (steps_done) *steps_done = done;
    return AKMC_SUCCESS;
}

int akmc_batch_kernel_timer(akmc_batch_t b, int enable)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    if (enable)
        for (auto &row : b->kev)
            for (auto &e : row)
                if (!e) HIPCHK(hipEventCreate(&e));
    b->ktimer = enable != 0;
    b->kms = 0.0;
    b->klaunches = 0;
    return AKMC_SUCCESS;
}

int akmc_batch_eval_counts(akmc_batch_t b, long long *counts)
{
    if (!b || !counts) return fail(AKMC_EINVAL, "bad arguments");
    if (!b->search_alloc) {
        for (int w = 0; w < b->W; ++w) counts[w] = 0;
        return AKMC_SUCCESS;
    }
    HIPCHK(hipMemcpy(counts, b->s.evals32, b->W * sizeof(long long), hipMemcpyDeviceToHost));
    return AKMC_SUCCESS;
}

int akmc_batch_kernel_time(akmc_batch_t b, double *ms, long long *launches)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    if (ms) *ms = b->kms;
    if (launches) *launches = b->klaunches;
    return AKMC_SUCCESS;
}

int akmc_batch_set_graphs(akmc_batch_t b, int enable)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    b->use_graphs = enable != 0;
    if (!b->use_graphs) drop_graph(b);
    return AKMC_SUCCESS;
}

int akmc_batch_set_logic_block(akmc_batch_t b, int threads)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    if (threads != 256 && threads != 1024) return fail(AKMC_EINVAL, "threads must be 256 or 1024, got %d", threads);
    b->logic_wide = threads == 1024;
    drop_graph(b);
    return AKMC_SUCCESS;
}

int akmc_batch_set_search_zone(akmc_batch_t b, int enable)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    b->zone_on = enable != 0;
    drop_graph(b);
    return AKMC_SUCCESS;
}

int akmc_batch_set_search_variant(akmc_batch_t b, const int v[9])
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    int nv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto ok = [](int x) { return x == 1 || x == 2 || x == 4 || x == 8; };
    if (v) {
        for (int p = 0; p < 3; ++p) {
            const int l = v[3 * p], u = v[3 * p + 1], st = v[3 * p + 2];
            if (l == 0) continue;
            if (!ok(l) || !ok(u)) return fail(AKMC_EINVAL, "part %d: lanes and unroll must be 1, 2, 4 or 8", p + 1);
            if (st && !b->stage_ok) return fail(AKMC_EUNSUPPORTED, "LDS staging is not possible for this batch");
            nv[3 * p] = l;
            nv[3 * p + 1] = u;
            nv[3 * p + 2] = st ? 1 : 0;
        }
    }
    std::memcpy(b->sv, nv, sizeof nv);
    drop_graph(b);
    return AKMC_SUCCESS;
}

static int time_search_phases_prec(akmc_batch_t b, int prec, int nrep, double ms[4]);

int akmc_batch_autotune_search(akmc_batch_t b, int nrep, int best[9], double ms1[AKMC_NVARIANTS],
                               double ms2[AKMC_NVARIANTS], double ms64[AKMC_NVARIANTS])
{
    if (!b || nrep <= 0) return fail(AKMC_EINVAL, "bad arguments");
    if (!b->search_alloc) return fail(AKMC_EINVAL, "no search has been started");
    static const int lanes[4] = {1, 2, 4, 8}, unroll[4] = {1, 2, 4, 8};
    const int nv = b->stage_ok ? AKMC_NVARIANTS : AKMC_NVARIANTS / 2;
    double t1[AKMC_NVARIANTS], t2[AKMC_NVARIANTS], ms[4];
    for (int v = 0; v < AKMC_NVARIANTS; ++v) {
        t1[v] = t2[v] = -1.0;
        if (v >= nv) continue;
        const int sv[9] = {lanes[v / 4 % 4], unroll[v % 4], v / 16, lanes[v / 4 % 4], unroll[v % 4], v / 16, 0, 0, 0};
        TRY(akmc_batch_set_search_variant(b, sv));
        TRY(akmc_batch_time_search_phases(b, 1, ms));  // first launch of this instantiation
        TRY(akmc_batch_time_search_phases(b, nrep, ms));
        t1[v] = ms[1];
        t2[v] = ms[2];
    }
    int k1 = 0, k2 = 0;
    for (int v = 1; v < nv; ++v) {
        if (t1[v] < t1[k1]) k1 = v;
        if (t2[v] < t2[k2]) k2 = v;
    }
    // the FP64 passes of mixed steps, on the walkers under FP64 verification
    // at the moment: both passes together
    double t64[AKMC_NVARIANTS];
    int k64 = 0;
    for (int v = 0; v < AKMC_NVARIANTS; ++v) {
        t64[v] = -1.0;
        if (v >= nv) continue;
        const int sv[9] = {0, 0, 0, 0, 0, 0, lanes[v / 4 % 4], unroll[v % 4], v / 16};
        TRY(akmc_batch_set_search_variant(b, sv));
        TRY(time_search_phases_prec(b, 2 | 4, 1, ms));
        TRY(time_search_phases_prec(b, 2 | 4, nrep, ms));
        t64[v] = ms[1] + ms[2];
        if (t64[v] < t64[k64]) k64 = v;
    }
    const int sv[9] = {lanes[k1 / 4 % 4], unroll[k1 % 4], k1 / 16, lanes[k2 / 4 % 4], unroll[k2 % 4], k2 / 16,
                       lanes[k64 / 4 % 4], unroll[k64 % 4], k64 / 16};
    TRY(akmc_batch_set_search_variant(b, sv));
    if (best) std::memcpy(best, sv, sizeof sv);
    for (int v = 0; v < AKMC_NVARIANTS; ++v) {
        if (ms1) ms1[v] = t1[v];
        if (ms2) ms2[v] = t2[v];
        if (ms64) ms64[v] = t64[v];
    }
    return AKMC_SUCCESS;
}

int akmc_batch_status(akmc_batch_t b, akmc_walker_status *out)
{
    if (!b || !out) return fail(AKMC_EINVAL, "bad arguments");
    if (!b->search_alloc) return fail(AKMC_EINVAL, "no search has been started");
    // while a run chunk is in flight: the states at the end of the previous
    // chunk, with the starts made since
    if (!b->inflight) TRY(fetch_states(b));
    for (int w = 0; w < b->W; ++w) {
        const WalkerState &ws = b->hws[w];
        out[w].kind = ws.kind;
        out[w].status = ws.status;
        out[w].done = ws.state == ST_DONE ? 1 : 0;
        out[w].iterations = ws.it;
        out[w].force_calls = ws.fcalls;
        out[w].energy = ws.E0;
        out[w].curvature = ws.C;
        out[w].verified = ws.verified_state;
        out[w].energy32 = ws.e32;
        out[w].curvature32 = ws.c32;
        out[w].energy64 = ws.e64;
        out[w].fmax64 = ws.fmax64;
        out[w].lambda1 = ws.lam1;
        out[w].lambda2 = ws.lam2;
        out[w].new_basin = ws.b2_new;
        out[w].basin2_status = ws.b2_status;
        out[w].basin2_force_calls = ws.b2_calls;
        out[w].basin2_iterations = ws.b2_iters;
        out[w].energy_basin2 = ws.eb2;
        out[w].energy_basin2_64 = ws.eb2_64;
        out[w].basin2_max_disp = std::sqrt(ws.b2_maxd2);
    }
    return AKMC_SUCCESS;
}

int akmc_batch_get_positions(akmc_batch_t b, double *x, double *y, double *z)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    size_t A = b->A;
    // before the first search or relax the batch's own positions are current
    const double *src[3] = {b->d.x, b->d.y, b->d.z};
    if (b->search_alloc) {
        src[0] = b->s.X;
        src[1] = b->s.X + A;
        src[2] = b->s.X + 2 * A;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    if (x) HIPCHK(hipMemcpy(x, src[0], A * sizeof(double), hipMemcpyDeviceToHost));
    if (y) HIPCHK(hipMemcpy(y, src[1], A * sizeof(double), hipMemcpyDeviceToHost));
    if (z) HIPCHK(hipMemcpy(z, src[2], A * sizeof(double), hipMemcpyDeviceToHost));
    return AKMC_SUCCESS;
}

int akmc_batch_get_walkers_positions(akmc_batch_t b, int nw, const int *walkers, double *out)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    if (b->search_alloc) return copy_walkers_out(b, b->s.X, nw, walkers, out);
    const double *const src[3] = {b->d.x, b->d.y, b->d.z};
    return copy_walkers_out3(b, src, nw, walkers, out);
}

int akmc_batch_get_walkers_saddles(akmc_batch_t b, int nw, const int *walkers, double *out)
{
    if (!b || !b->search_alloc || !b->s.Xs) return fail(AKMC_EINVAL, "the basin-2 stage has not been set up");
    return copy_walkers_out(b, b->s.Xs, nw, walkers, out);
}

}  // extern "C"

namespace {

// the [3][n] slices of src ([3][A]) of the listed walkers, one pinned
// transfer
int copy_walkers_out(akmc_batch_t b, const double *src, int nw, const int *walkers, double *out)
{
    const double *const c3[3] = {src, src + b->A, src + 2 * (size_t)b->A};
    return copy_walkers_out3(b, c3, nw, walkers, out);
}

// the flagged walkers' (3, n_w) slices of three [A] component arrays, packed
// walker after walker, through the pinned staging buffer
int copy_walkers_out3(akmc_batch_t b, const double *const src[3], int nw, const int *walkers, double *out)
{
    if (nw < 0 || (nw && (!walkers || !out))) return fail(AKMC_EINVAL, "bad arguments");
    size_t total = 0;
    for (int k = 0; k < nw; ++k) {
        if (walkers[k] < 0 || walkers[k] >= b->W) return fail(AKMC_EINVAL, "walker %d out of range", walkers[k]);
        total += 3 * (size_t)b->hw[walkers[k]].n;
    }
    if (nw == 0) return AKMC_SUCCESS;
    // one buffer: walker ids, their output offsets, then the packed doubles
    const size_t ibytes = ((size_t)nw * sizeof(int) + 7) / 8 * 8, obytes = (size_t)nw * sizeof(long long);
    const size_t need = ibytes + obytes + total * sizeof(double);
    if (need > b->gat_n) {
        if (b->gat_pin) HIPCHK(hipHostFree(b->gat_pin));
        if (b->gat_dev) HIPCHK(hipFree(b->gat_dev));
        b->gat_pin = b->gat_dev = nullptr;
        b->gat_n = 0;
        const size_t cap = need + need / 2;
        HIPCHK(hipHostMalloc((void **)&b->gat_pin, cap, hipHostMallocDefault));
        HIPCHK(hipMalloc((void **)&b->gat_dev, cap));
        b->gat_n = cap;
    }
    // while a run chunk is in flight the copy runs beside it on a side
    // stream (the walkers that had ended before the chunk are not written)
    hipStream_t st = b->stream;
    if (b->inflight) {
        if (!b->side) HIPCHK(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
        st = b->side;
    }
    int *ws = reinterpret_cast<int *>(b->gat_pin);
    long long *dst = reinterpret_cast<long long *>(b->gat_pin + ibytes);
    long long o = 0;
    for (int k = 0; k < nw; ++k) {
        ws[k] = walkers[k];
        dst[k] = o;
        o += 3 * (long long)b->hw[walkers[k]].n;
    }
    HIPCHK(hipMemcpyAsync(b->gat_dev, b->gat_pin, ibytes + obytes, hipMemcpyHostToDevice, st));
    double *dout = reinterpret_cast<double *>(b->gat_dev + ibytes + obytes);
    akmc_launch_gather_walkers(b->d.walkers, src, reinterpret_cast<const int *>(b->gat_dev),
                               reinterpret_cast<const long long *>(b->gat_dev + ibytes), nw, dout, st);
    HIPCHK(hipGetLastError());
    double *hout = reinterpret_cast<double *>(b->gat_pin + ibytes + obytes);
    HIPCHK(hipMemcpyAsync(hout, dout, total * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::memcpy(out, hout, total * sizeof(double));
    return AKMC_SUCCESS;
}

}  // namespace

extern "C" {

int akmc_batch_get_walker_positions(akmc_batch_t b, int walker, double *x, double *y, double *z)
{
    if (!b || walker < 0 || walker >= b->W) return fail(AKMC_EINVAL, "bad arguments");
    size_t A = b->A, off = b->hw[walker].off, n = b->hw[walker].n;
    double *dst[3] = {x, y, z};
    const double *src[3] = {b->d.x, b->d.y, b->d.z};  // before the first search or relax
    if (b->search_alloc)
        for (int c = 0; c < 3; ++c) src[c] = b->s.X + c * A;
    for (int c = 0; c < 3; ++c)
        if (dst[c]) HIPCHK(hipMemcpyAsync(dst[c], src[c] + off, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}

int akmc_batch_get_mode(akmc_batch_t b, double *mode)
{
    if (!b || !b->search_alloc || !mode) return fail(AKMC_EINVAL, "no search state");
    HIPCHK(hipMemcpy(mode, b->s.N, 3 * (size_t)b->A * sizeof(double), hipMemcpyDeviceToHost));
    return AKMC_SUCCESS;
}


int akmc_batch_time_phases(akmc_batch_t b, int nrep, double ms[4])
{
    if (!b || nrep <= 0 || !ms) return fail(AKMC_EINVAL, "bad arguments");
    for (auto &e : b->evp)
        if (!e) HIPCHK(hipEventCreate(&e));
    for (int ph = 0; ph < 4; ++ph) ms[ph] = 0.0;
    for (int r = 0; r < nrep; ++r) {
        HIPCHK(hipEventRecord(b->evp[0], b->stream));
        for (int ph = 0; ph < 4; ++ph) {
            akmc_launch_phase(b->d, b->f->host, batch_mode(b), 0, ph, b->stream);
            HIPCHK(hipEventRecord(b->evp[ph + 1], b->stream));
        }
        HIPCHK(hipEventSynchronize(b->evp[4]));
        for (int ph = 0; ph < 4; ++ph) {
            float t = 0.0f;
            HIPCHK(hipEventElapsedTime(&t, b->evp[ph], b->evp[ph + 1]));
            ms[ph] += t;
        }
    }
    return check_overflow(b);
}

int akmc_batch_time_search_phases(akmc_batch_t b, int nrep, double ms[4])
{
    if (!b || nrep <= 0 || !ms) return fail(AKMC_EINVAL, "bad arguments");
    if (!b->search_alloc) return fail(AKMC_EINVAL, "no search has been started");
    return time_search_phases_prec(b, batch_mode(b), nrep, ms);
}

// per-phase times of nrep evaluations of the search walkers of precision
// mask prec (2 | 4: the FP64 passes of mixed steps) at their probes
static int time_search_phases_prec(akmc_batch_t b, int prec, int nrep, double ms[4])
{
    for (auto &e : b->evp)
        if (!e) HIPCHK(hipEventCreate(&e));
    const BatchDev d = step_dev(b);
    for (int ph = 0; ph < 4; ++ph) ms[ph] = 0.0;
    for (int r = 0; r < nrep; ++r) {
        HIPCHK(hipEventRecord(b->evp[0], b->stream));
        for (int ph = 0; ph < 4; ++ph) {
            akmc_launch_phase(d, b->f->host, prec, 0, ph, b->stream);
            HIPCHK(hipEventRecord(b->evp[ph + 1], b->stream));
        }
        HIPCHK(hipEventSynchronize(b->evp[4]));
        for (int ph = 0; ph < 4; ++ph) {
            float t = 0.0f;
            HIPCHK(hipEventElapsedTime(&t, b->evp[ph], b->evp[ph + 1]));
            ms[ph] += t;
        }
    }
    return check_overflow(b);
}

int akmc_batch_get_zone(akmc_batch_t b, int *member, int *size, int *flags, int *active)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    if (!b->search_alloc) return fail(AKMC_EINVAL, "no search has been started");
    if (member) HIPCHK(hipMemcpyAsync(member, b->zmark, b->A * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    if (size) HIPCHK(hipMemcpyAsync(size, b->nz, b->W * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    if (flags) HIPCHK(hipMemcpyAsync(flags, b->zflag, b->W * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    if (active) HIPCHK(hipMemcpyAsync(active, b->d.active, b->W * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}

int akmc_batch_save_basin(akmc_batch_t b) { return akmc_batch_set_basins(b, nullptr, nullptr); }

int akmc_batch_set_basins(akmc_batch_t b, const int *walkers, const double *x)
{
    if (!b) return fail(AKMC_EINVAL, "bad arguments");
    TRY(search_alloc(b));  // search positions start at the resident ones
    if (!b->basin) TRY(dalloc(b, &b->basin, 3 * (size_t)b->A));
    b->s.basin = b->basin;
    const size_t A = b->A;
    if (!walkers && !x) {
        HIPCHK(hipMemcpyAsync(b->basin, b->s.X, 3 * A * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    } else {
        for (int w = 0; w < b->W; ++w) {
            if (walkers && !walkers[w]) continue;
            for (int c = 0; c < 3; ++c) {
                const size_t o = c * A + b->hw[w].off;
                if (x)
                    HIPCHK(hipMemcpyAsync(b->basin + o, x + o, b->hw[w].n * sizeof(double), hipMemcpyHostToDevice,
                                          b->stream));
                else
                    HIPCHK(hipMemcpyAsync(b->basin + o, b->s.X + o, b->hw[w].n * sizeof(double),
                                          hipMemcpyDeviceToDevice, b->stream));
            }
        }
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}

void akmc_basin2_defaults(akmc_basin2_params *p)
{
    if (!p) return;
    p->basin2Delta = 0.2;
    p->kmcBasinRadius = 0.5;  // docs/examples/gdc/control:11
    p->kmcMaxRadius = 5.0;
}

int akmc_batch_set_basin2(akmc_batch_t b, const akmc_basin2_params *p)
{
    if (!b) return fail(AKMC_EINVAL, "null batch");
    TRY(search_alloc(b));
    if (!p) {
        b->s.basin2_on = 0;
        b->s.final64 = 0;
        return AKMC_SUCCESS;
    }
    if (!(p->basin2Delta >= 0.0) || !(p->kmcBasinRadius >= 0.0) || !(p->kmcMaxRadius >= 0.0))
        return fail(AKMC_EINVAL, "basin2Delta, kmcBasinRadius and kmcMaxRadius must be >= 0");
    if (!b->s.Xs) TRY(dalloc(b, &b->s.Xs, 3 * (size_t)b->A));
    b->s.basin2_delta = p->basin2Delta;
    b->s.basin_r2 = p->kmcBasinRadius * p->kmcBasinRadius;
    b->s.max_r2 = p->kmcMaxRadius * p->kmcMaxRadius;
    b->s.basin2_on = 1;
    b->s.final64 = 1;
    return AKMC_SUCCESS;
}

int akmc_batch_start_relax_at(akmc_batch_t b, const int *walkers, const double *x0, const int *frozen,
                              const akmc_relax_params *p)
{
    if (!b || !p) return fail(AKMC_EINVAL, "bad arguments");
    TRY(search_alloc(b));
    const size_t A = b->A;
    for (int w = 0; w < b->W; ++w) {
        if (walkers && !walkers[w]) continue;
        const int off = b->hw[w].off, n = b->hw[w].n;
        if (frozen) {
            std::copy(frozen + off, frozen + off + n, b->hfrozen.begin() + off);
            HIPCHK(hipMemcpyAsync((void *)(b->s.frozen + off), frozen + off, n * sizeof(int), hipMemcpyHostToDevice,
                                  b->stream));
        }
        if (x0)
            for (int c = 0; c < 3; ++c)
                HIPCHK(hipMemcpyAsync(b->s.X + c * A + off, x0 + c * A + off, n * sizeof(double),
                                      hipMemcpyHostToDevice, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    if (frozen) TRY(upload_active(b, walkers));
    return akmc_batch_start_relax(b, walkers, p);
}

int akmc_batch_cancel(akmc_batch_t b, const int *walkers)
{
    if (!b || !b->search_alloc) return fail(AKMC_EINVAL, "no search state");
    TRY(fetch_states(b));
    for (int w = 0; w < b->W; ++w)
        if (!walkers || walkers[w]) b->hws[w].kind = WK_IDLE;
    HIPCHK(hipMemcpyAsync(b->s.ws, b->hws.data(), b->W * sizeof(WalkerState), hipMemcpyHostToDevice, b->stream));
    int run;
    return refresh_active(b, &run);
}

int akmc_batch_restore_basin(akmc_batch_t b, const int *walkers)
{
    if (!b || !b->basin) return fail(AKMC_EINVAL, "no saved basin");
    size_t A = b->A;
    for (int w = 0; w < b->W; ++w) {
        if (walkers && !walkers[w]) continue;
        for (int c = 0; c < 3; ++c) {
            size_t o = c * A + b->hw[w].off;
            HIPCHK(hipMemcpyAsync(b->s.X + o, b->basin + o, b->hw[w].n * sizeof(double), hipMemcpyDeviceToDevice,
                                  b->stream));
        }
        if (!walkers) break;
    }
    if (!walkers)
        HIPCHK(hipMemcpyAsync(b->s.X, b->basin, 3 * A * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return AKMC_SUCCESS;
}


int akmc_batch_set_verify(akmc_batch_t b, int nl, double delta)
{
    if (!b || nl < 0 || nl > AKMC_LZ_MAX || (nl > 0 && delta <= 0.0))
        return fail(AKMC_EINVAL, "verification needs 0 <= nl <= %d and delta > 0", AKMC_LZ_MAX);
    TRY(search_alloc(b));
    if (nl > 0) TRY(ensure_lanczos(b));
    if (nl > 0 && (size_t)nl > (size_t)AKMC_LZ_MAX) return fail(AKMC_EINVAL, "nl too large");
    b->s.verify_nl = nl;
    b->s.verify_delta = delta;
    drop_graph(b);
    return AKMC_SUCCESS;
}

int akmc_batch_search_energies(akmc_batch_t b, int bits, double *energies)
{
    if (!b || !b->search_alloc || !energies || (bits != 32 && bits != 64))
        return fail(AKMC_EINVAL, "bad arguments");
    size_t A = b->A;
    HIPCHK(hipMemcpyAsync(b->d.x, b->s.X, A * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d.y, b->s.X + A, A * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d.z, b->s.X + 2 * A, A * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    std::vector<int> mode(b->W, bits == 64 ? 2 : 1);
    HIPCHK(hipMemcpyAsync(b->d.active, mode.data(), b->W * sizeof(int), hipMemcpyHostToDevice, b->stream));
    akmc_launch_forces_impl(b->d, b->f->host, bits == 64 ? 2 : 1, 0, b->stream);
    HIPCHK(hipGetLastError());
    TRY(check_overflow(b));
    TRY(fetch_out(b));
    for (int w = 0; w < b->W; ++w) energies[w] = b->hout[w].total;
    return AKMC_SUCCESS;
}

}  // extern "C"

// ------------------------------------------------------------------ MLP

struct akmc_mlp_s {
    int rows = 0;
    unsigned seed = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *w1t = nullptr, *w2t = nullptr, *w3t = nullptr, *w2r = nullptr, *w3r = nullptr;
    float *X = nullptr, *T = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Y = nullptr,
          *dY = nullptr, *dZ2 = nullptr, *dZ1 = nullptr, *P = nullptr;
    double *lpart = nullptr;
    int adam_t = 0;
    static constexpr int kChunk = 256;
};

namespace {
template <typename T>
int malloc_m(akmc_mlp_t m, T **p, size_t n)
{
    HIPCHK(hipMalloc((void **)p, (n ? n : 1) * sizeof(T)));
    m->allocs.push_back((void *)*p);
    return AKMC_SUCCESS;
}

void mlp_free(akmc_mlp_t m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *q : m->allocs) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int mlp_refresh_weights(akmc_mlp_t m)
{
    akmc_mlp_launch_weights(m->p, m->w1t, m->w2t, m->w3t, m->w2r, m->w3r, m->stream);
    HIPCHK(hipGetLastError());
    return AKMC_SUCCESS;
}

// forward + loss/gradient of rows [0, M) already in m->X / m->T
int mlp_loss_grad_dev(akmc_mlp_t m, int M, double *loss)
{
    akmc_mlp_launch_forward(m->X, M, m->p, m->w1t, m->w2t, m->w3t, m->Z1, m->H1, m->Z2, m->H2, m->Y, m->stream);
    akmc_mlp_launch_backward(m->X, m->T, M, m->w2r, m->w3r, m->Z1, m->H1, m->Z2, m->H2, m->Y, m->dY, m->dZ2,
                             m->dZ1, m->lpart, m->P, akmc_mlp_s::kChunk, m->g, m->stream);
    HIPCHK(hipGetLastError());
    if (loss) {
        int nb = (M + 255) / 256;
        std::vector<double> part(nb);
        HIPCHK(hipMemcpyAsync(part.data(), m->lpart, nb * sizeof(double), hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        double s = 0.0;
        for (double v : part) s += v;
        *loss = s / M;
    }
    return AKMC_SUCCESS;
}
}  // namespace

extern "C" {

int akmc_mlp_create(int max_rows, unsigned seed, akmc_mlp_t *out)
{
    if (max_rows <= 0 || !out) return fail(AKMC_EINVAL, "bad arguments");
    akmc_mlp_t m = new akmc_mlp_s();
    m->rows = (max_rows + 255) / 256 * 256;
    m->seed = seed;
    auto bad = [&](int rc) {
        mlp_free(m);
        return rc;
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return bad(fail(AKMC_EHIP, "stream creation failed"));
    const size_t R = m->rows, NP = AKMC_MLP_NPARAM;
    int rc = AKMC_SUCCESS;
    if ((rc = malloc_m(m, &m->p, NP)) || (rc = malloc_m(m, &m->g, NP)) || (rc = malloc_m(m, &m->m1, NP)) ||
        (rc = malloc_m(m, &m->m2, NP)) || (rc = malloc_m(m, &m->w1t, 128 * 48)) ||
        (rc = malloc_m(m, &m->w2t, 128 * 128)) || (rc = malloc_m(m, &m->w3t, 16 * 128)) ||
        (rc = malloc_m(m, &m->w2r, 128 * 128)) || (rc = malloc_m(m, &m->w3r, 128 * 16)) ||
        (rc = malloc_m(m, &m->X, R * 48)) || (rc = malloc_m(m, &m->T, R * 3)) ||
        (rc = malloc_m(m, &m->Z1, R * 128)) || (rc = malloc_m(m, &m->H1, R * 128)) ||
        (rc = malloc_m(m, &m->Z2, R * 128)) || (rc = malloc_m(m, &m->H2, R * 128)) ||
        (rc = malloc_m(m, &m->Y, R * 16)) || (rc = malloc_m(m, &m->dY, R * 16)) ||
        (rc = malloc_m(m, &m->dZ2, R * 128)) || (rc = malloc_m(m, &m->dZ1, R * 128)) ||
        (rc = malloc_m(m, &m->P, (R / akmc_mlp_s::kChunk + 1) * 128 * 128)) ||
        (rc = malloc_m(m, &m->lpart, R / 256 + 1)))
        return bad(rc);
    HIPCHK(hipMemset(m->m1, 0, NP * sizeof(float)));
    HIPCHK(hipMemset(m->m2, 0, NP * sizeof(float)));
    // He-style initial weights from a seeded generator, zero biases
    std::vector<float> h(NP, 0.0f);
    unsigned long long st = 0x9E3779B97F4A7C15ull ^ seed;
    auto unif = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto gauss = [&]() { return std::sqrt(-2.0 * std::log(unif())) * std::cos(6.283185307179586 * unif()); };
    size_t o = 0;
    for (int i = 0; i < 48 * 128; ++i) h[o++] = (float)(gauss() * std::sqrt(2.0 / 48));
    o += 128;
    for (int i = 0; i < 128 * 128; ++i) h[o++] = (float)(gauss() * std::sqrt(2.0 / 128));
    o += 128;
    for (int i = 0; i < 128 * 3; ++i) h[o++] = (float)(gauss() * std::sqrt(1.0 / 128));
    HIPCHK(hipMemcpy(m->p, h.data(), NP * sizeof(float), hipMemcpyHostToDevice));
    if ((rc = mlp_refresh_weights(m))) return bad(rc);
    HIPCHK(hipStreamSynchronize(m->stream));
    *out = m;
    return AKMC_SUCCESS;
}

void akmc_mlp_destroy(akmc_mlp_t m) { mlp_free(m); }

int akmc_mlp_set_params(akmc_mlp_t m, const float *params)
{
    if (!m || !params) return fail(AKMC_EINVAL, "bad arguments");
    HIPCHK(hipMemcpyAsync(m->p, params, AKMC_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
    TRY(mlp_refresh_weights(m));
    HIPCHK(hipStreamSynchronize(m->stream));
    return AKMC_SUCCESS;
}

int akmc_mlp_get_params(akmc_mlp_t m, float *params)
{
    if (!m || !params) return fail(AKMC_EINVAL, "bad arguments");
    HIPCHK(hipMemcpyAsync(params, m->p, AKMC_MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return AKMC_SUCCESS;
}

int akmc_mlp_forward(akmc_mlp_t m, const float *X, int M, float *Y)
{
    if (!m || !X || !Y || M < 0) return fail(AKMC_EINVAL, "bad arguments");
    std::vector<float> y16;
    for (int r0 = 0; r0 < M; r0 += m->rows) {
        int n = std::min(m->rows, M - r0);
        HIPCHK(hipMemcpyAsync(m->X, X + (size_t)r0 * 48, (size_t)n * 48 * sizeof(float), hipMemcpyHostToDevice,
                              m->stream));
        akmc_mlp_launch_forward(m->X, n, m->p, m->w1t, m->w2t, m->w3t, m->Z1, m->H1, m->Z2, m->H2, m->Y, m->stream);
        HIPCHK(hipGetLastError());
        y16.resize((size_t)n * 16);
        HIPCHK(hipMemcpyAsync(y16.data(), m->Y, (size_t)n * 16 * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) Y[(size_t)(r0 + i) * 3 + c] = y16[(size_t)i * 16 + c];
    }
    return AKMC_SUCCESS;
}

int akmc_mlp_loss_grad(akmc_mlp_t m, const float *X, const float *T, int M, double *loss, float *grad)
{
    if (!m || !X || !T || M <= 0 || M > m->rows) return fail(AKMC_EINVAL, "M must be in [1, max_rows]");
    HIPCHK(hipMemcpyAsync(m->X, X, (size_t)M * 48 * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->T, T, (size_t)M * 3 * sizeof(float), hipMemcpyHostToDevice, m->stream));
    TRY(mlp_loss_grad_dev(m, M, loss));
    if (grad) {
        HIPCHK(hipMemcpyAsync(grad, m->g, AKMC_MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
    }
    return AKMC_SUCCESS;
}

int akmc_mlp_time(akmc_mlp_t m, int M, int nrep, double ms[2])
{
    if (!m || !ms || M <= 0 || M > m->rows || nrep <= 0) return fail(AKMC_EINVAL, "M must be in [1, max_rows]");
    HIPCHK(hipMemsetAsync(m->X, 0, (size_t)M * 48 * sizeof(float), m->stream));
    HIPCHK(hipMemsetAsync(m->T, 0, (size_t)M * 3 * sizeof(float), m->stream));
    hipEvent_t e[3];
    for (auto &v : e) HIPCHK(hipEventCreate(&v));
    TRY(mlp_loss_grad_dev(m, M, nullptr));  // first launches outside the timing
    HIPCHK(hipEventRecord(e[0], m->stream));
    for (int r = 0; r < nrep; ++r)
        akmc_mlp_launch_forward(m->X, M, m->p, m->w1t, m->w2t, m->w3t, m->Z1, m->H1, m->Z2, m->H2, m->Y, m->stream);
    HIPCHK(hipEventRecord(e[1], m->stream));
    for (int r = 0; r < nrep; ++r) TRY(mlp_loss_grad_dev(m, M, nullptr));
    HIPCHK(hipEventRecord(e[2], m->stream));
    HIPCHK(hipEventSynchronize(e[2]));
    float t0 = 0.0f, t1 = 0.0f;
    HIPCHK(hipEventElapsedTime(&t0, e[0], e[1]));
    HIPCHK(hipEventElapsedTime(&t1, e[1], e[2]));
    for (auto &v : e) (void)hipEventDestroy(v);
    ms[0] = t0 / nrep;
    ms[1] = t1 / nrep;
    return AKMC_SUCCESS;
}

int akmc_mlp_train(akmc_mlp_t m, const float *X, const float *T, int M, int epochs, int batch, float lr,
                   unsigned seed, double *loss_history)
{
    if (!m || !X || !T || M <= 0 || epochs <= 0 || batch <= 0 || batch > m->rows)
        return fail(AKMC_EINVAL, "bad training arguments (batch must be <= max_rows)");
    float *dX = nullptr, *dT = nullptr;
    int *didx = nullptr;
    double *dacc = nullptr;  // per-epoch loss accumulator (no host sync per minibatch)
    HIPCHK(hipMalloc(&dX, (size_t)M * 48 * sizeof(float)));
    HIPCHK(hipMalloc(&dT, (size_t)M * 3 * sizeof(float)));
    HIPCHK(hipMalloc(&didx, (size_t)M * sizeof(int)));
    HIPCHK(hipMalloc(&dacc, (size_t)epochs * sizeof(double)));
    HIPCHK(hipMemset(dacc, 0, (size_t)epochs * sizeof(double)));
    auto cleanup = [&]() {
        (void)hipFree(dX);
        (void)hipFree(dT);
        (void)hipFree(didx);
        (void)hipFree(dacc);
    };
    if (hipMemcpy(dX, X, (size_t)M * 48 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dT, T, (size_t)M * 3 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return fail(AKMC_EHIP, "training data upload failed");
    }
    std::vector<int> perm(M);
    unsigned long long st = 0xD1B54A32D192ED03ull ^ seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    int nbatch_ep = 1;
    for (int ep = 0; ep < epochs; ++ep) {
        // the previous epoch's copy of perm must be done before perm changes
        if (ep > 0 && hipStreamSynchronize(m->stream) != hipSuccess) {
            cleanup();
            return fail(AKMC_EHIP, "training step failed");
        }
        for (int i = 0; i < M; ++i) perm[i] = i;
        for (int i = M - 1; i > 0; --i) std::swap(perm[i], perm[next() % (unsigned long long)(i + 1)]);
        if (hipMemcpyAsync(didx, perm.data(), (size_t)M * sizeof(int), hipMemcpyHostToDevice, m->stream) !=
            hipSuccess) {
            cleanup();
            return fail(AKMC_EHIP, "permutation upload failed");
        }
        int nb = 0;
        for (int b0 = 0; b0 + batch <= M || (b0 == 0 && M < batch); b0 += batch) {
            int n = std::min(batch, M - b0);
            akmc_mlp_launch_gather(dX, 48, didx + b0, n, m->X, m->stream);
            akmc_mlp_launch_gather(dT, 3, didx + b0, n, m->T, m->stream);
            int rc = mlp_loss_grad_dev(m, n, nullptr);
            if (rc != AKMC_SUCCESS) {
                cleanup();
                return rc;
            }
            akmc_mlp_launch_loss_accum(m->lpart, (n + 255) / 256, 1.0 / n, dacc + ep, m->stream);
            akmc_mlp_launch_adam(m->p, m->g, m->m1, m->m2, AKMC_MLP_NPARAM, lr, ++m->adam_t, m->stream);
            rc = mlp_refresh_weights(m);
            if (rc != AKMC_SUCCESS) {
                cleanup();
                return rc;
            }
            ++nb;
            if (M < batch) break;
        }
        nbatch_ep = std::max(nb, 1);
    }
    std::vector<double> acc(epochs);
    hipError_t e = hipMemcpyAsync(acc.data(), dacc, (size_t)epochs * sizeof(double), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    cleanup();
    if (e != hipSuccess) return fail(AKMC_EHIP, "training failed: %s", hipGetErrorString(e));
    if (loss_history)
        for (int ep = 0; ep < epochs; ++ep) loss_history[ep] = acc[ep] / nbatch_ep;
    return AKMC_SUCCESS;
}

int akmc_batch_descriptors(akmc_batch_t b, int nq, const int *atom, const double *vac, const double *mover,
                           float *X)
{
    if (!b || nq <= 0 || !atom || !vac || !mover || !X) return fail(AKMC_EINVAL, "bad arguments");
    for (int q = 0; q < nq; ++q)
        if (atom[q] < 0 || atom[q] >= b->A) return fail(AKMC_EINVAL, "query %d: atom out of range", q);
    int *qa;
    double *qv, *qm;
    float *qx;
    HIPCHK(hipMalloc(&qa, nq * sizeof(int)));
    HIPCHK(hipMalloc(&qv, 3 * (size_t)nq * sizeof(double)));
    HIPCHK(hipMalloc(&qm, 3 * (size_t)nq * sizeof(double)));
    HIPCHK(hipMalloc(&qx, 48 * (size_t)nq * sizeof(float)));
    HIPCHK(hipMemcpyAsync(qa, atom, nq * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(qv, vac, 3 * (size_t)nq * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(qm, mover, 3 * (size_t)nq * sizeof(double), hipMemcpyHostToDevice, b->stream));
    akmc_mlp_launch_descriptors(b->d, nq, qa, qv, qm, qx, b->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(X, qx, 48 * (size_t)nq * sizeof(float), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(qa);
    (void)hipFree(qv);
    (void)hipFree(qm);
    (void)hipFree(qx);
    if (e != hipSuccess) return fail(AKMC_EHIP, "descriptors: %s", hipGetErrorString(e));
    return AKMC_SUCCESS;
}

}  // extern "C"
