/*
 * akmc_oracle.h -- CPU FP64 oracle for the force-and-saddle-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * arithmetic (jap93/AKMC "ACDC"), used by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg as the checker and the timed CPU
 * baseline.  The product library (akmc_amd/, libakmc_hip.so) never links,
 * loads or calls anything here.
 *
 * Conventions (same as the reference's Field classes):
 *   - positions, forces: SoA FP64 arrays x[n], y[n], z[n];
 *   - lat[9]: lattice vectors as rows (a1 = lat[0..2]);
 *   - rcp[9]: inverse of lat, indexed so that the fractional coordinate
 *     s_a = sum_b r_b * rcp[3*b + a]  (potentials/metal/Field.cpp:396-398);
 *   - stress[6] = xx, yy, zz, yz, xz, xy.
 *
 * Each function cites the reference lines it follows.  Parity of every
 * function is pinned by tests/test_oracle_*.py (closed forms, invariants,
 * finite differences, brute force, reference-held data); see the header
 * comment of each test for which check pins what.
 */
#ifndef AKMC_ORACLE_H
#define AKMC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define AKO_MAXSPEC 8
#define AKO_MAXPOT 36

/* Physical constants fixed by this build (the reference's Constants.h is
 * absent from its tree).  e^2/(4 pi eps0) in eV*A as 14.3997584 (the value that reproduces the
 * reference-held energy of docs/examples/gdc/basis:3 to 1e-9 relative) and
 * the eV -> 10 J/mol factor that scales FIRE/ART time steps
 * (potentials/metal/Energy.h:40), DL_POLY's 9648.530821 (trajectory only;
 * no reference data fixes it). */
#define AKO_CTOEV 14.3997584
#define AKO_EVTODL 9648.530821

/* status codes: Common/Status.h:11 (TERMINATED's value is this build's) */
#define AKO_SUCCESS 0
#define AKO_FAILED 1
#define AKO_INTERRUPTED 3
#define AKO_TERMINATED 4

/* ------------------------------------------------------------------ */
/* cell helpers                                                        */
/* ------------------------------------------------------------------ */
void ako_dcell(const double lat[9], double out[10]);
void ako_inverse_cell(const double lat[9], double rcp[9]);
void ako_wrap_min_image(int n, double *x, double *y, double *z,
                        const double lat[9], const double rcp[9]);
void ako_wrap_no_image(int n, double *x, double *y, double *z,
                       const double lat[9], const double rcp[9]);

/* ------------------------------------------------------------------ */
/* Sutton-Chen many-body field (potentials/metal/)                     */
/* par: 9 doubles per potential: eps, a, n, m, c (ManyBody.cpp:151-163) */
/* table[nspec*nspec]: potential index per species pair, -1 = none     */
/* ------------------------------------------------------------------ */
typedef struct {
    int nspec;
    int npot;
    int table[AKO_MAXSPEC * AKO_MAXSPEC];
    double par[AKO_MAXPOT * 9];
    double rc;
    int min_image;        /* 1: minimum image; 0: explicit images */
} ako_sc_field;

/* returns total energy; e_many/e_pair/rho may be NULL */
double ako_sc_forces(const ako_sc_field *f, int n, double *x, double *y, double *z,
                     const int *type, const double lat[9], const double rcp[9],
                     double *fx, double *fy, double *fz, double stress[6],
                     double *e_many, double *e_pair, double *rho);
/* same arithmetic, OpenMP over the outer atom loop (Field.cpp:293-490) */
double ako_sc_forces_omp(const ako_sc_field *f, int n, double *x, double *y, double *z,
                         const int *type, const double lat[9], const double rcp[9],
                         double *fx, double *fy, double *fz, double stress[6],
                         double *e_many, double *e_pair);
double ako_sc_energy(const ako_sc_field *f, int n, double *x, double *y, double *z,
                     const int *type, const double lat[9], const double rcp[9],
                     double *e_many, double *e_pair);
/* number of ordered in-cutoff pairs under the field's image convention */
long long ako_sc_count_pairs(const ako_sc_field *f, int n, const double *x, const double *y,
                             const double *z, const double lat[9], const double rcp[9]);

/* ------------------------------------------------------------------ */
/* Ewald ionic field with short-range pair forms (potentials/pair/)    */
/* par: 6 doubles per potential, in the order the potentials text      */
/* gives them (TwoBody.cpp:98-245): buck A rho C; morse D alpha r0 C12; */
/* ljones eps sigma; bhm A B sigma C D; severnsix eps sigma            */
/* ------------------------------------------------------------------ */
#define AKO_BUCK 0
#define AKO_MORSE 1
#define AKO_LJONES 2
#define AKO_BHM 3
#define AKO_SEVNSIX 4
typedef struct {
    int nspec;
    int npot;
    int table[AKO_MAXSPEC * AKO_MAXSPEC];
    double par[AKO_MAXPOT * 6];
    double rc;
    int min_image;
    int do_ewald;         /* 1: Ewald, 0: none */
    double eta;
    double cfact;         /* CTOEV */
    int cells[3];         /* explicit image shells for min_image == 0 */
    int nk;               /* number of k vectors */
    double *gx, *gy, *gz, *g0, *g1;
    int form[AKO_MAXPOT]; /* AKO_BUCK ... per potential */
    int energy_path;      /* 1: short-range energies of TwoBody's energy-only form */
    /* do_ewald == 2: damped-shifted Coulomb on the erfc mesh of
     * ako_damped_mesh (Field::fieldErfcGen, pair Field.cpp:19-73) */
    const double *dsf_e, *dsf_f;
    double dsf_del, dsf_rfld0, dsf_rfld1;
} ako_ion_field;

/* Field::fieldErfcGen (potentials/pair/Field.cpp:19-73): nmesh points
 * (the reference's MAXMESH) of erfc(alpha r)/r by Abramowitz-Stegun 7.1.26
 * and of its force term, spacing rc / (nmesh - 4); returns the spacing */
double ako_damped_mesh(double alpha, double rc, int nmesh, double *e, double *f, double *rfld0, double *rfld1);

void ako_ewald_eta_from_acc(double acc, double rc, double *eta, double *eps, double *tol);
void ako_ewald_gcell(const double lat[9], const double rcp[9], double eta, double rc,
                     double eps, double tol, int gcell[3], double *rcpcut);
int ako_ewald_kvectors(const double rcp[9], double volume, double eta, double cfact,
                       const int gcell[3], double rcpcut, int maxk,
                       double *gx, double *gy, double *gz, double *g0, double *g1);
double ako_ewald_recip_force(int n, const double *x, const double *y, const double *z,
                             const double *q, int nk, const double *gx, const double *gy,
                             const double *gz, const double *g0, const double *g1,
                             double *fx, double *fy, double *fz, double stress[6]);
double ako_ewald_recip_energy(int n, const double *x, const double *y, const double *z,
                              const double *q, int nk, const double *gx, const double *gy,
                              const double *gz, const double *g0);
/* full driver (pair Field::calculateForces, Field.cpp:323-381).
 * returns total; e_rcp/e_real/e_vdw may be NULL */
double ako_ion_forces(const ako_ion_field *f, int n, double *x, double *y, double *z,
                      const int *type, const double *q, const double lat[9], const double rcp[9],
                      double *fx, double *fy, double *fz, double stress[6],
                      double *e_rcp, double *e_real, double *e_vdw);
long long ako_ion_count_pairs(const ako_ion_field *f, int n, const double *x, const double *y,
                              const double *z, const double lat[9], const double rcp[9]);

/* ------------------------------------------------------------------ */
/* generic force callback used by the optimisers                       */
/* must overwrite f*, may wrap x,y,z in place; returns total energy     */
/* ------------------------------------------------------------------ */
typedef double (*ako_force_fn)(void *ctx, int n, double *x, double *y, double *z,
                               double *fx, double *fy, double *fz);

typedef struct {
    const void *field;
    int kind;             /* 0 = SC, 1 = ionic, 2 = SC threaded (OpenMP) */
    const int *type;
    const double *q;
    const double *lat;
    const double *rcp;
    long long ncalls;
} ako_field_ctx;
double ako_field_cb(void *ctx, int n, double *x, double *y, double *z,
                    double *fx, double *fy, double *fz);

/* ------------------------------------------------------------------ */
/* FIRE relax (optim/Relax.cpp:44-315), RelaxCntrl (Common/RelaxCntrl.h) */
/* ------------------------------------------------------------------ */
typedef struct {
    int maxIter;
    double normTol, forceTol;
    int N_min;
    double initTimeStep, alphaDecrease, stepIncrease, stepDecrease, alphaStart,
           maxTimeStep;
    int euler;            /* 1 = semi-implicit Euler (reference default) */
    double timeStepUnit;  /* EVTODL */
    /* FIRE2 (Relax::FIRE2Relax) */
    int N_delaySteps, N_maxUphill, initialDelay;
    double minTimeStep;
} ako_relax_cntrl;
void ako_relax_defaults(ako_relax_cntrl *c);
int ako_fire_relax(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                   const int *frozen, const ako_relax_cntrl *c, double *energy, int *iters);
/* Relax::FIRE2Relax (optim/Relax.cpp:317-648) */
int ako_fire2_relax(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                    const int *frozen, const ako_relax_cntrl *c, double *energy, int *iters);

/* ------------------------------------------------------------------ */
/* saddle-search helpers (optim/SaddleSearch.cpp)                      */
/* ------------------------------------------------------------------ */
int ako_active_sphere(int n, const double *x, const double *y, const double *z,
                      const double lat[9], const double rcp[9], int centre, double radius,
                      int *frozen);
/* u: 3*nactive uniforms in [0,1); writes unit rvec and displaces atoms */
void ako_random_displacement(int n, double *x, double *y, double *z, const int *frozen,
                             const double *u, double d0, double *rvec);

/* tridiagonal symmetric eigensolver (replaces LAPACK dstev, ART.cpp:764-794):
 * d[n] diagonal, e[n-1] sub-diagonal; returns 0 on success, eigenvalues
 * ascending in w[n], eigenvectors as columns of z[n*n] (column-major) */
int ako_tridiag_eig(int n, const double *d, const double *e, double *w, double *z);

/* ------------------------------------------------------------------ */
/* ART (optim/ART.cpp:17-756), ARTCntrl (Common/ArtCntrl.h)            */
/* ------------------------------------------------------------------ */
typedef struct {
    int numLanczosVectors;
    double deltaX, initialDisplacement, maxEigenvalue, eigValTol, parDamp;
    int goPerpendicular;
    int maxParIter, maxPerpIter;
    double normTol, forceTol;
    int N_min;
    double initTimeStep, alphaDecrease, stepIncrease, stepDecrease, alphaStart,
           maxTimeStep, initTimeStepPerp, maxTimeStepPerp;
    double timeStepUnit;
} ako_art_cntrl;
void ako_art_defaults(ako_art_cntrl *c);
/* positions are the already displaced start; rvec (3*nactive) the unit
 * start direction.  On return x,y,z hold the saddle. */
int ako_art_search(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                   const int *frozen, double *rvec, const ako_art_cntrl *c,
                   double *saddle_energy, double *eigval, int *iters);
/* one Lanczos lowest-mode evaluation (ART.cpp:133-371) */
int ako_lanczos(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                const int *frozen, double *rvec, int nlanczos, double deltaX, double eigValTol,
                double *eigval, double *fpar, double *fperp, double *energy);

/* ------------------------------------------------------------------ */
/* Dimer (no live reference code: Henkelman & Jonsson 1999, Olsen et   */
/* al. 2004, following optim/Dimer.cpp:416-619).  DimerCntrl names.    */
/* ------------------------------------------------------------------ */
typedef struct {
    double deltaTheta, dimerDisplacement, maxDeltaX;
    int maxRotations;
    int maxParIter;
    double maxEigenvalue, normTol, forceTol;
    int N_min;
    double alphaDecrease, stepIncrease, stepDecrease, alphaStart;
    double initTimeStep, maxTimeStep, maxStep, timeStepUnit;
} ako_dimer_cntrl;
void ako_dimer_defaults(ako_dimer_cntrl *c);
int ako_dimer_search(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                     const int *frozen, double *mode, const ako_dimer_cntrl *c,
                     double *saddle_energy, double *curvature, int *iters, int *fcalls);

#ifdef __cplusplus
}
#endif
#endif
