/*
 * akmc.h -- C ABI of libakmc_hip.so, the MI355X engine for AKMC's
 * force-and-saddle-search path.
 *
 * Plain pointers and sizes only.  Positions and forces are SoA FP64 arrays
 * x[n], y[n], z[n] as in the reference's Field::calculateForces; lattices
 * are 9 doubles, lattice vectors as rows; stress is xx,yy,zz,yz,xz,xy.
 * Every entry point returns an int status: AKMC_SUCCESS (0), AKMC_FAILED
 * (1), AKMC_INTERRUPTED (3), AKMC_TERMINATED (4) as in Common/Status.h:10-50,
 * or a negative library error whose text akmc_last_error() returns.  The
 * library never calls exit().  Handles are not re-entrant: one host thread
 * drives one handle.
 */
#ifndef AKMC_H
#define AKMC_H

#ifdef __cplusplus
extern "C" {
#endif

#define AKMC_SUCCESS 0
#define AKMC_FAILED 1
#define AKMC_INTERRUPTED 3
#define AKMC_TERMINATED 4
#define AKMC_EINVAL (-1)      /* bad argument or input text */
#define AKMC_EHIP (-2)        /* HIP runtime error */
#define AKMC_ECAPACITY (-3)   /* a fixed device capacity was exceeded */
#define AKMC_EUNSUPPORTED (-4)

/* Energy record: the fields of Energy (potentials/metal/Energy.h:17-37,
 * potentials/pair/Energy.h); total = rcp + real + vdw + many + ext
 * (Energy::getTotalEnergy, potentials/metal/Energy.h:58-59). */
typedef struct {
    double total;
    double rcp;
    double real;
    double many;
    double vdw;
    double ext;
} akmc_energy;

typedef struct akmc_field_s *akmc_field_t;
typedef struct akmc_batch_s *akmc_batch_t;

const char *akmc_last_error(void);
int akmc_version(void);

/* ---- force field ("plugin") ------------------------------------------
 * Replaces Field::readPotential (potentials/metal/Field.cpp:122-205,
 * potentials/pair/Field.cpp:215-317): the same "potentials" text, i.e.
 * cutoff / species / manybody|many + suttonchen / twobody + buck /
 * ewald precis|none / noimage / close.  kind is inferred from the text. */
int akmc_field_create(const char *potentials_text, akmc_field_t *out);
void akmc_field_destroy(akmc_field_t f);
/* 0 = many-body (Sutton-Chen), 1 = ionic (Buckingham + Ewald) */
int akmc_field_kind(akmc_field_t f);
/* Explicit Ewald parameters (eta, |n|max per axis) instead of the
 * accuracy-derived ones of potentials/pair/Field.cpp:129-132 and
 * Ewald.cpp:88-120; call before akmc_field_setup. */
int akmc_field_set_ewald(akmc_field_t f, double eta, int gx, int gy, int gz);
/* Verlet skin of the device neighbour lists (default 0.7 A). */
int akmc_field_set_skin(akmc_field_t f, double skin);
/* Number density (atoms per A^3) the neighbour lists of later batches are
 * sized for, when it exceeds a walker's own N / V: walkers that hold a
 * local cluster of a larger crystal in that crystal's periodic box (0, the
 * default: each walker's N / V).  No reference counterpart (the reference
 * keeps no neighbour lists on this path). */
int akmc_field_set_list_density(akmc_field_t f, double atoms_per_A3);
/* Pair-physics precision of later force calls: 32 (FP32 pair terms, the
 * default) or 64 (FP64 throughout; the verification path).  Cutoff tests,
 * minimum images, reciprocal sums and all accumulations are FP64 in both. */
int akmc_field_set_precision(akmc_field_t f, int bits);
/* Replaces Field::setup (potentials/metal/Field.cpp:24-97,
 * potentials/pair/Field.cpp:80-169): cutoff checks, image shells,
 * k vectors.  Orthorhombic cells only on the device. */
int akmc_field_setup(akmc_field_t f, const double lat[9], int max_atoms);
/* Replaces Field::calculateForces (potentials/metal/Field.cpp:211-276,
 * potentials/pair/Field.cpp:323-381): x,y,z are wrapped in place as the
 * reference does; f* and stress are overwritten. */
int akmc_field_forces(akmc_field_t f, double *x, double *y, double *z, double *fx, double *fy,
                      double *fz, double stress[6], const double lat[9], const double *charge,
                      const int *type, const int *frozen, int n, akmc_energy *e);
/* Replaces Field::calculateEnergy (energy only). */
int akmc_field_energy(akmc_field_t f, double *x, double *y, double *z, const double lat[9],
                      const double *charge, const int *type, int n, akmc_energy *e);
/* number of k vectors of the last setup (0 for many-body fields) */
int akmc_field_num_kvectors(akmc_field_t f);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  One Sutton-Chen force evaluation of the configuration
 * in the caller's device arrays x, y, z [n] (FP64, not wrapped) and type
 * [n], with the kernels and default variant of akmc_field_forces, so the
 * same results bit for bit: forces into the device arrays fx, fy, fz [n],
 * the energy record into the device akmc_energy *energy (may be NULL) and
 * the overflow bits (0: the list capacity sufficed) into the device int
 * *flags (may be NULL).  Enqueued on stream (a hipStream_t; NULL: the null
 * stream); returns without waiting.  Allocates nothing: every other buffer
 * lives in one fixed device array of the library, akmc_noalloc_scratch_bytes()
 * per GPU, held until the process ends; calls that overlap on one GPU race
 * on it.  The field must have been set up (akmc_field_setup) and be a
 * Sutton-Chen field (AKMC_EUNSUPPORTED otherwise); n above
 * AKMC_NOALLOC_MAX_ATOMS is AKMC_ECAPACITY; neighbour lists hold up to
 * AKMC_NOALLOC_MAX_NEIGHBOURS entries per atom. */
#define AKMC_NOALLOC_MAX_ATOMS 8192
#define AKMC_NOALLOC_MAX_NEIGHBOURS 320
int akmc_field_forces_noalloc(akmc_field_t f, const double *x, const double *y, const double *z,
                              const int *type, int n, const double lat[9], double *fx, double *fy,
                              double *fz, akmc_energy *energy, int *flags, void *stream);
unsigned long akmc_noalloc_scratch_bytes(void);

/* ---- device-resident walker batches ----------------------------------
 * W independent cells of uneven size kept in HBM; the batched counterpart
 * of the reference's MPI task farm (kmc/KineticMonteCarlo.cpp:352-598),
 * one walker per search. */
int akmc_batch_create(akmc_field_t f, int nwalkers, const int *natoms, const double *lat,
                      akmc_batch_t *out);
void akmc_batch_destroy(akmc_batch_t b);
/* all walkers' atoms concatenated in walker order; charge may be NULL */
int akmc_batch_upload(akmc_batch_t b, const double *x, const double *y, const double *z,
                      const int *type, const double *charge, const int *frozen);
int akmc_batch_download(akmc_batch_t b, double *x, double *y, double *z, double *fx, double *fy,
                        double *fz);
/* one force evaluation of every walker at its resident positions;
 * energies[W] (may be NULL) receives each walker's energy record */
int akmc_batch_forces(akmc_batch_t b, akmc_energy *energies);
/* device-resident repeat: nrep force evaluations with no host copies,
 * returns the total device time in ms measured with HIP events */
int akmc_batch_forces_repeat(akmc_batch_t b, int nrep, double *ms);
/* per-phase device time of nrep force evaluations at the resident
 * positions, HIP events on the batch's stream: ms[0] list maintenance,
 * ms[1] pass 1 (Sutton-Chen density / ionic real space), ms[2] pass 2
 * (Sutton-Chen force / Ewald reciprocal), ms[3] per-walker reduction;
 * totals over the nrep evaluations */
int akmc_batch_time_phases(akmc_batch_t b, int nrep, double ms[4]);
/* Autotuner of the Sutton-Chen pair kernels.  A kernel variant is
 * (lanes, unroll, staged): each atom's neighbour list is split over `lanes`
 * consecutive lanes of a wavefront (1, 2, 4 or 8), combined by cross-lane
 * sums, and every lane keeps `unroll` neighbour gathers in flight (1, 2, 4
 * or 8).  Small walkers fill the chip better with more lanes per atom; more
 * gathers in flight hide memory latency at the cost of registers.  Staged
 * variants run one 1024-thread workgroup per chunk of 1024 / lanes atoms of
 * one walker, which first copies the walker's packed positions into LDS, so
 * the neighbour gathers read LDS instead of L2 (possible when the largest
 * walker's 32 bytes per atom fit one workgroup's LDS; MI355X: 160 KiB, about
 * 4900 atoms).  akmc_batch_set_variant and akmc_batch_set_staging fix the
 * choice (default 1, 1, unstaged; the graph of a running search is
 * recaptured); akmc_batch_staging reports the current mode and whether
 * staging is possible (AKMC_EUNSUPPORTED from set_staging otherwise).
 * akmc_batch_autotune times nrep force evaluations at the resident positions
 * for each variant, ms[16*s + 4*l + u] for staged s (0, 1), lanes (1, 2, 4,
 * 8)[l] and unroll (1, 2, 4, 8)[u] (-1 where staging is not possible), keeps
 * the fastest and returns it in best[0] (lanes), best[1] (unroll), best[2]
 * (staged).  Variants with the same lanes agree bitwise in FP32 (to
 * FMA-contraction round-off in FP64), staged or not; different lanes agree
 * to summation-order round-off.  In a micro-step that mixes FP32 walkers
 * with walkers under FP64 verification, the FP64 kernels run unstaged (they
 * serve few walkers, which finer workgroups spread over more CUs).
 * No reference counterpart: the reference evaluates one configuration per
 * call on the host (potentials/metal/Field.cpp:293-490). */
#define AKMC_NVARIANTS 32
int akmc_batch_set_variant(akmc_batch_t b, int lanes, int unroll);
int akmc_batch_set_staging(akmc_batch_t b, int on);
int akmc_batch_staging(akmc_batch_t b, int *on, int *possible);
int akmc_batch_autotune(akmc_batch_t b, int nrep, int best[3], double ms[AKMC_NVARIANTS]);
int akmc_batch_total_atoms(akmc_batch_t b);
/* neighbour-list rebuild count since creation (diagnostics) */
long long akmc_batch_rebuilds(akmc_batch_t b);

/* ---- searches on device-resident walkers ------------------------------
 * Each walker advances by "micro-steps": one batched force evaluation at
 * the walker's probe positions followed by one update of its state
 * machine.  All walkers of a batch move together; finished walkers are
 * skipped by the force kernels. */

/* RelaxCntrl (Common/RelaxCntrl.h:15-48).  Relax::RelaxStructure
 * (optim/Relax.cpp:19-42) dispatches on method: AKMC_RELAX_FIRE is
 * Relax::FIRERelax (:44-315), AKMC_RELAX_FIRE2 Relax::FIRE2Relax
 * (:317-648), AKMC_RELAX_NONE one force and energy evaluation reported as
 * success (calculateConfig, :33-39).  verletIntegration != 0 selects the
 * velocity-Verlet branches of both (:194-262, :519-585) instead of the
 * semi-implicit Euler default.  N_delaySteps, N_maxUphill, initialDelay and
 * minTimeStep are read by FIRE2 only.  timeStepUnit is the eV -> 10 J/mol
 * factor EVTODL (potentials/metal/Energy.h:40). */
#define AKMC_RELAX_FIRE 0
#define AKMC_RELAX_FIRE2 1
#define AKMC_RELAX_NONE 2
typedef struct {
    int maxIter;
    double normTol, forceTol;
    int N_min;
    double initTimeStep, alphaDecrease, stepIncrease, stepDecrease, alphaStart, maxTimeStep;
    double timeStepUnit;
    int method;
    int verletIntegration;
    int N_delaySteps, N_maxUphill, initialDelay;
    double minTimeStep;
} akmc_relax_params;

/* DimerCntrl (Common/DimerCntrl.h:18-88) plus the translation time steps
 * and per-atom step cap this build adds (the reference Dimer is a stub,
 * optim/Dimer.cpp:15-42). */
typedef struct {
    double deltaTheta, dimerDisplacement, maxDeltaX;
    int maxRotations, maxParIter;
    double maxEigenvalue, normTol, forceTol;
    int N_min;
    double alphaDecrease, stepIncrease, stepDecrease, alphaStart;
    double initTimeStep, maxTimeStep, maxStep, timeStepUnit;
} akmc_dimer_params;

/* ARTCntrl (Common/ArtCntrl.h:14-74); the perpendicular pre-phase
 * (goPeprepedicular) is not offered on the device */
typedef struct {
    int numLanczosVectors;
    double deltaX, maxEigenvalue, eigValTol, parDamp;
    int maxParIter;
    double normTol, forceTol;
    int N_min;
    double initTimeStep, alphaDecrease, stepIncrease, stepDecrease, alphaStart, maxTimeStep;
    double timeStepUnit;
} akmc_art_params;

typedef struct {
    int kind;        /* 0 idle, 1 relax, 2 dimer, 3 FP64 verification, 4 ART, 5 basin-2 relax */
    int status;      /* AKMC_SUCCESS / AKMC_FAILED once done */
    int done;
    int iterations;
    int force_calls;
    double energy;   /* energy at the current centre */
    double curvature;
    /* FP64 verification of a converged dimer (akmc_batch_set_verify) */
    int verified;        /* 1 once the FP64 data below is present */
    double energy32;     /* search energy at the saddle (FP32 pair terms) */
    double curvature32;  /* dimer curvature at convergence */
    double energy64;     /* FP64 energy at the saddle */
    double fmax64;       /* largest FP64 atomic force on the free atoms */
    double lambda1, lambda2; /* two lowest FP64 Lanczos curvatures */
    /* basin 2 of a converged search (akmc_batch_set_basin2); with the stage
     * on, a relax walker's energy64 is its FP64 energy at the minimum */
    int new_basin;            /* 1: Basis::checkBasin accepted basin 2 */
    int basin2_status;        /* AKMC_SUCCESS / AKMC_FAILED of its FIRE relax */
    int basin2_force_calls;
    int basin2_iterations;
    double energy_basin2;     /* at basin 2, search precision */
    double energy_basin2_64;  /* at basin 2, FP64 pair terms */
    double basin2_max_disp;   /* largest minimum-image displacement from the basin (A) */
} akmc_walker_status;

/* Basin 2 of every converged search: the part of
 * KineticMonteCarlo::checkForNewBasin (kmc/KineticMonteCarlo.cpp:1626-1713)
 * that follows findSaddlePoint.  Names of the reference's kmc{} keys
 * basin2delta, kmcbasinradius (docs/docs/source/ACDCUserGuide.rst:150-180)
 * and of the KMCCntrl members read at kmc/KineticMonteCarlo.cpp:1672,1690. */
typedef struct {
    double basin2Delta;     /* kick: free atoms move by basin2Delta |D| D from the saddle */
    double kmcBasinRadius;  /* a new basin has an atom displaced by at least this (A) */
    double kmcMaxRadius;    /* ... and none by more than this (A) */
} akmc_basin2_params;

void akmc_relax_defaults(akmc_relax_params *p);
void akmc_dimer_defaults(akmc_dimer_params *p);
void akmc_art_defaults(akmc_art_params *p);
/* frozen[A] (1 = fixed), as Basis::frozen (Common/Basis.cpp:902-981) */
int akmc_batch_set_frozen(akmc_batch_t b, const int *frozen);
/* Relax::RelaxStructure (optim/Relax.h:116-117) for the walkers with
 * walkers[w] != 0 (all when walkers is NULL), from their current positions */
int akmc_batch_start_relax(akmc_batch_t b, const int *walkers, const akmc_relax_params *p);
/* Dimer::findSaddlePoint (optim/Dimer.h:68-70) for the walkers with
 * walkers[w] != 0 (all when NULL).  x0 ([3][A], may be NULL to keep the
 * current positions) holds the displaced start, mode ([3][A]) the start
 * direction, frozen ([A], may be NULL to keep the current mask) the fixed
 * atoms; only the flagged walkers' slices of the arrays are read. */
int akmc_batch_start_dimer(akmc_batch_t b, const int *walkers, const double *x0, const double *mode,
                           const int *frozen, const akmc_dimer_params *p);
/* ART::findSaddlePoint (optim/ART.cpp:17-129) without the perpendicular
 * pre-phase: x0/mode/frozen as for the dimer, mode is the initial rVector */
int akmc_batch_start_art(akmc_batch_t b, const int *walkers, const double *x0, const double *mode,
                         const int *frozen, const akmc_art_params *p);
/* run micro-steps until every walker is done or max_steps have run; the
 * host looks at the walker states every sync_every steps (1 = the plain,
 * host-driven loop); steps of sync_every > 1 replay a captured hipGraph */
int akmc_batch_run(akmc_batch_t b, int max_steps, int sync_every, int *steps_done);
/* the two halves of one akmc_batch_run chunk, so the host can work while the
 * device runs it (no counterpart in the reference): run_async enqueues
 * steps micro-steps and the status copy and returns; run_wait waits for
 * them and reports the steps run (0 when no walker was running) and the
 * walkers still running.  Between the two, the only calls allowed on the
 * batch are akmc_batch_status (which then returns the states at the end of
 * the previous chunk, with the starts made since), the walker position and
 * saddle copies (made beside the chunk: defined for walkers that had ended
 * before it) and akmc_batch_start_dimer / akmc_batch_start_art of walkers
 * that had ended (queued behind the chunk: they run from the next one). */
int akmc_batch_run_async(akmc_batch_t b, int steps);
int akmc_batch_run_wait(akmc_batch_t b, int *steps_done, int *running);
/* akmc_batch_run_async without its host round trip (the states read back
 * and the active flags written before the launch): for a caller that only
 * changes walkers through the akmc_batch_start_* calls between chunks, so
 * the device's active flags and the host's copy of the states (the last
 * wait's, with the starts made since) are current.  A pipelined caller
 * launches the next chunk right after a wait and reads and refills the
 * ended searches while it runs.  No counterpart in the reference. */
int akmc_batch_run_next(akmc_batch_t b, int steps);
/* harvest (no counterpart in the reference): while on, every run chunk ends
 * by copying the search positions of the walkers that ended in it (and of
 * those that had ended before harvest was turned on) into one of two
 * mapped host buffers, used in turn, so a pipelined caller reads them while
 * the next chunk runs.  akmc_batch_get_harvested_positions returns, laid
 * out as akmc_batch_get_walkers_positions, those of walkers harvested by
 * the chunk of the last akmc_batch_run_wait (other walkers' slices hold
 * older data).  Needs a started search; not while a chunk is in flight. */
int akmc_batch_set_harvest(akmc_batch_t b, int on);
int akmc_batch_get_harvested_positions(akmc_batch_t b, int nw, const int *walkers, double *out);
/* enable (default) or disable the hipGraph replay of akmc_batch_run; when
 * disabled the same kernels are launched one by one on the batch's stream
 * (for tools that cannot trace graph launches).  Results are identical. */
int akmc_batch_set_graphs(akmc_batch_t b, int enable);
/* enable (default) or disable the active zones of searches on minimum-image
 * Sutton-Chen fields (no counterpart in the reference, whose force call
 * evaluates every atom).  A search moves only its free atoms, so its micro-
 * steps evaluate forces on the free atoms and densities and energies on the
 * atoms within the list radius of a free atom; every other atom keeps the
 * values of the walker's first full evaluation since its start, which no
 * free atom can change.  The searches' forces on free atoms, energies and
 * trajectories are bit for bit those of full evaluations; forces on frozen
 * atoms are not re-evaluated during a search (the optimizers zero them), so
 * akmc_batch_download after a search returns their values from its start:
 * call akmc_batch_forces first for fresh ones. */
int akmc_batch_set_search_zone(akmc_batch_t b, int enable);
/* Workgroup size of the walker state machines: 256 threads, or 1024 (the
 * default for batches whose largest walker has at least 16 384 atoms).  No
 * counterpart in the reference; the two differ only in the order of their
 * block sums. */
int akmc_batch_set_logic_block(akmc_batch_t b, int threads);
/* akmc_batch_time_phases for the search walkers at their current probes, as
 * the search micro-steps launch the kernels (active zones included), without
 * advancing any search (no counterpart in the reference: measurement). */
int akmc_batch_time_search_phases(akmc_batch_t b, int nrep, double ms[4]);
/* The active zones (no counterpart in the reference: inspection): member[A]
 * 1 for the atoms in their walker's zone, size[W] zone sizes, flags[W] the
 * zone bits (1 on, 2 / 4 a full FP32 / FP64 evaluation pending, 8 reset
 * pending), active[W] the walkers' evaluation modes (0 none, 1 FP32 pair
 * terms, 2 FP64); any pointer may be NULL. */
int akmc_batch_get_zone(akmc_batch_t b, int *member, int *size, int *flags, int *active);
/* Kernel variants of the search micro-steps' Sutton-Chen passes (no
 * counterpart in the reference): v = {lanes, unroll, staged} of pass 1
 * (density; it also writes the pair energies in search micro-steps), of
 * pass 2 (forces), then of both FP64 passes of steps that mix FP32 and FP64
 * walkers; lanes 0 keeps the default for that part, NULL resets all.
 * Results do not depend on unroll or staging; lanes change only the
 * summation order of the sums they split. */
int akmc_batch_set_search_variant(akmc_batch_t b, const int v[9]);
/* Times every variant of each pass on the search walkers at their current
 * probes (akmc_batch_time_search_phases, nrep evaluations each; the FP64
 * variants on the walkers under FP64 verification), keeps the fastest per
 * part and returns it in best[9] (may be NULL) with the per-variant ms of
 * pass 1, pass 2 and the FP64 passes, indexed like akmc_batch_autotune (any
 * of them NULL).  Call it with searches under way. */
int akmc_batch_autotune_search(akmc_batch_t b, int nrep, int best[9], double ms1[AKMC_NVARIANTS],
                               double ms2[AKMC_NVARIANTS], double ms64[AKMC_NVARIANTS]);
/* Kernel timer of akmc_batch_run: HIP events on the batch's stream around
 * the FP32 Sutton-Chen pass-2 launch (k_sc_force) of every micro-step.
 * Events inside a replayed graph cannot be timed, so while the timer runs
 * the micro-steps are launched one by one.  akmc_batch_kernel_timer(b, 1)
 * clears and starts it (sync_every <= 64 while it runs), (b, 0) stops it;
 * akmc_batch_kernel_time returns the accumulated milliseconds and the
 * number of timed launches.  No reference counterpart (measurement). */
int akmc_batch_kernel_timer(akmc_batch_t b, int enable);
int akmc_batch_kernel_time(akmc_batch_t b, double *ms, long long *launches);
/* per walker: FP32 force evaluations the walker's state machines have
 * consumed since the search state was allocated (counts[W]) */
int akmc_batch_eval_counts(akmc_batch_t b, long long *counts);
int akmc_batch_status(akmc_batch_t b, akmc_walker_status *out);
/* current search positions (the dimer centre / relaxed positions); before
 * the batch's first search or relax, the positions it was created or last
 * loaded with (akmc_batch_get_walker(s)_positions alike) */
int akmc_batch_get_positions(akmc_batch_t b, double *x, double *y, double *z);
int akmc_batch_get_mode(akmc_batch_t b, double *mode);
/* search positions of one walker (x, y, z of its n atoms) */
int akmc_batch_get_walker_positions(akmc_batch_t b, int walker, double *x, double *y, double *z);
/* search positions of several walkers in one transfer: out holds, for each
 * listed walker in order, its x[n], y[n], z[n] */
int akmc_batch_get_walkers_positions(akmc_batch_t b, int nw, const int *walkers, double *out);
/* keep the current search positions of every walker as its basin, and put
 * the walkers flagged in walkers (all when NULL) back on their basin */
int akmc_batch_save_basin(akmc_batch_t b);
/* FP64 re-verification of every converged dimer inside the same walker
 * stream: nl Lanczos steps (<= 32) of FP64 finite-difference Hessian
 * products with step delta, started on the dimer mode; nl = 0 disables */
int akmc_batch_set_verify(akmc_batch_t b, int nl, double delta);
/* energies of every walker at its search positions, FP32 (bits = 32) or
 * FP64 (bits = 64) pair terms */
int akmc_batch_search_energies(akmc_batch_t b, int bits, double *energies);
int akmc_batch_restore_basin(akmc_batch_t b, const int *walkers);

/* ---- KMC step on the device ------------------------------------------
 * checkForNewBasin (kmc/KineticMonteCarlo.cpp:1626-1713) after every
 * converged search, in the same walker stream (after its FP64
 * verification when that is on): the saddle is kept, the free atoms are
 * kicked by basin2Delta |D| D (D = minimum-image saddle - basin over all
 * atoms), FIRE-relaxed with the search's frozen mask (Relax::RelaxStructure
 * at :1686) and compared with the basin (Basis::checkBasin,
 * Common/Basis.cpp:731-781).  The basin is the one set by
 * akmc_batch_save_basin / akmc_batch_set_basins.  While the stage is on,
 * every relax (basin and basin 2) ends with one FP64 energy evaluation.
 * p = NULL turns the stage off. */
void akmc_basin2_defaults(akmc_basin2_params *p);
int akmc_batch_set_basin2(akmc_batch_t b, const akmc_basin2_params *p);
/* the saddle positions kept by the basin-2 stage, laid out as in
 * akmc_batch_get_walkers_positions (which then returns basin 2) */
int akmc_batch_get_walkers_saddles(akmc_batch_t b, int nw, const int *walkers, double *out);
/* set the basin of the walkers flagged in walkers (all when NULL) to x
 * ([3][A] host array, only the flagged slices are read) or, when x is
 * NULL, to their current search positions */
int akmc_batch_set_basins(akmc_batch_t b, const int *walkers, const double *x);
/* Relax::RelaxStructure (optim/Relax.h:116-117) of the flagged walkers
 * from x0 ([3][A], NULL: current positions) with the frozen mask
 * ([A], NULL: current mask); only the flagged walkers' slices are read.
 * The KMC loop relaxes each new basin 1 this way
 * (kmc/KineticMonteCarlo.cpp:110-115). */
int akmc_batch_start_relax_at(akmc_batch_t b, const int *walkers, const double *x0, const int *frozen,
                              const akmc_relax_params *p);
/* stop the flagged walkers (they become idle and are not evaluated), the
 * TERMINATED command the reference sends to search groups once enough
 * events are in (kmc/KineticMonteCarlo.cpp:446-455) */
int akmc_batch_cancel(akmc_batch_t b, const int *walkers);

/* ---- KMC step bookkeeping on the host (no device work) ----------------
 * Energies in eV, temperatures in K, rates in the prefactor's unit.  The
 * Boltzmann constant lives in the reference's missing Constants.h; this is
 * the CODATA 2018 value. */
#define AKMC_BOLTZMANN 8.617333262e-5
/* Event rate prefactor exp(-Ea / (kB T)) (kmc/Event.cpp:66-105) */
double akmc_kmc_rate(double activation_energy, double prefactor, double temperature);
/* largest accepted activation energy minCap + window kB T (kmc/kmc.cpp:160) */
double akmc_kmc_max_cap(double min_cap, double window, double temperature);
/* the event chosen by u in [0, 1): the first i whose cumulative rate
 * exceeds u * total (kmc/KineticMonteCarlo.cpp:144-169); total_rate may be
 * NULL; -1 when no event qualifies */
int akmc_kmc_select(int n, const double *rates, double u, double *total_rate);
/* KMC time step -(1/R) ln u, 0 when |R| <= 1e-18 (calculateTime,
 * kmc/KineticMonteCarlo.cpp:1151-1161, negated at :171) */
double akmc_kmc_time_increase(double total_rate, double u);
/* Basis::checkBasin (Common/Basis.cpp:731-781) for n atoms, positions SoA
 * [3][n], any cell (rows of lat are the lattice vectors): 1 when some atom
 * moved by >= radius and none by > max_radius */
int akmc_kmc_check_basin(int n, const double *basin1, const double *basin2, const double lat[9], double radius,
                         double max_radius);
/* Event::getDisplacedAtoms (kmc/Event.cpp:303-341): the atoms of pos
 * displaced from basin by more than radius, written to atoms[]; returns
 * their count */
int akmc_kmc_displaced_atoms(int n, const double *pos, const double *basin, const double lat[9], double radius,
                             int *atoms);
/* clearEventList (kmc/KineticMonteCarlo.cpp:1163-1265): keep[i] = 1 for
 * the events that are recycled after event `chosen` ran: recycle on, centre
 * at least radius from the chosen centre (minimum image), for mode 2 a
 * draw u[i] below percent / 100, and blocked[i] == 0 (blocked: the event
 * displaces an atom of the excluded list; NULL when superBasin is off).
 * centres[3 * i + k]; returns the number kept */
int akmc_kmc_recycle(int n, const double *centres, int chosen, const double lat[9], double radius, int recycle,
                     int mode, double percent, const double *u, const int *blocked, int *keep);
/* updateExcludedList (kmc/KineticMonteCarlo.cpp:1291-1313): count every
 * positive entry of excluded[natoms] down by one, then add history to the
 * displaced atoms */
int akmc_kmc_update_excluded(int natoms, int *excluded, int nd, const int *displaced, int history);

/* The KMC event packet of sendEvent / receiveEvent
 * (kmc/KineticMonteCarlo.cpp:1376-1461) and Event(int*, float*, natoms,
 * prefactor, temperature) (kmc/Event.cpp:66-105): veci[n] the atoms'
 * global numbers; vecf[6n + 7] = per atom basin-2 x y z, saddle x y z, then
 * basin-1, basin-2, saddle and activation energies and the centre, all
 * FP32.  Unpacking returns FP64 copies of the FP32 values, energies[4] in
 * that order, and the event's rate. */
int akmc_kmc_event_pack(int n, const int *global_no, const double *basin2, const double *saddle, double e1,
                        double e2, double es, double ea, const double centre[3], int *veci, float *vecf);
int akmc_kmc_event_unpack(int n, const int *veci, const float *vecf, double prefactor, double temperature,
                          int *global_no, double *basin2, double *saddle, double energies[4], double centre[3],
                          double *rate);

/* ---- text formats (host) ----------------------------------------------
 * species: the species names separated by blanks; an atom's label indexes
 * it.  Positions are SoA [3][n].  The writers use the reference's iostream
 * formats, byte for byte. */
/* Basis::dumpBasis (Common/Basis.cpp:600-649): the title and "0 3 N"
 * header when header != 0 (the reference writes it on an object's first
 * dump only), then "timeStep it N 0 3 time energy", the lattice and the
 * atoms; append != 0 adds to the file (the archive) */
int akmc_io_dump_basis(const char *path, int append, int header, const char *title, const char *species, int n,
                       const int *label, const double *pos, const double lat[9], double simulation_time,
                       double energy, int iteration);
/* runKMC's restart file (kmc/KineticMonteCarlo.cpp:210-218): "restart
 * title", " 0 3 N N", then dumpBasis */
int akmc_io_write_restart(const char *path, int header, const char *title, const char *species, int n,
                          const int *label, const double *pos, const double lat[9], double simulation_time,
                          double energy, int iteration);
/* Basis::inputBasis (Common/Basis.cpp:356-501): akmc_io_basis_size reads the
 * atom count and levcfg; akmc_io_read_basis the labels, positions, lattice,
 * the restart iteration and time (0 when absent) and the title */
int akmc_io_basis_size(const char *path, int *n, int *levcfg);
int akmc_io_read_basis(const char *path, const char *species, int n, int *label, double *pos, double lat[9],
                       int *iteration, double *time, char *title, int title_len);
/* writeEventList (kmc/KineticMonteCarlo.cpp:1561-1616) with
 * Event::writeCentre / writeSaddle / writeBasin2 (kmc/Event.cpp:216-292):
 * basin 1, then per event its activation energy, centre[3], saddle energy and
 * positions (saddles [nev][3][n]) and basin-2 energy and positions.
 * fp32_events != 0 prints the event quantities as the reference holds them,
 * rounded to FP32 by its event packet. */
int akmc_io_write_events(const char *path, const char *species, int n, const int *label, const double *basin1,
                         const double lat[9], double basin1_energy, int nev, const double *act_energy,
                         const double *centres, const double *saddle_energy, const double *saddles,
                         const double *basin2_energy, const double *basin2s, int chosen, double simulation_time,
                         int fp32_events);

/* ---- learned proposal (no counterpart in the reference) ---------------
 * Per-atom descriptors (32 radial + 16 angular symmetry functions with
 * vacancy-site and candidate-mover channels) and the MLP 48 -> 128 -> 128
 * -> 3 with SiLU, BF16 MFMA GEMMs with FP32 accumulation, FP32 master
 * weights.  Parameters are 23171 floats in the order W1[48][128], b1[128],
 * W2[128][128], b2[128], W3[128][3], b3[3] (row-major). */
#define AKMC_MLP_NPARAM 23171
typedef struct akmc_mlp_s *akmc_mlp_t;
int akmc_mlp_create(int max_rows, unsigned seed, akmc_mlp_t *out);
void akmc_mlp_destroy(akmc_mlp_t m);
int akmc_mlp_set_params(akmc_mlp_t m, const float *params);
int akmc_mlp_get_params(akmc_mlp_t m, float *params);
/* Y[M][3] = MLP(X[M][48]) (host arrays, M <= max_rows per call chunk) */
int akmc_mlp_forward(akmc_mlp_t m, const float *X, int M, float *Y);
/* mean-squared loss against T[M][3] and its gradient (layout of params) */
int akmc_mlp_loss_grad(akmc_mlp_t m, const float *X, const float *T, int M, double *loss, float *grad);
/* Adam training on device-resident data: epochs passes in seeded random
 * minibatches; loss_history[epochs] receives each epoch's mean loss */
/* device time of one forward (ms[0]) and of one forward + backward
 * (loss gradient, ms[1]) over M <= max_rows zero rows, HIP events on the
 * network's stream, averaged over nrep launches (measurement) */
int akmc_mlp_time(akmc_mlp_t m, int M, int nrep, double ms[2]);
int akmc_mlp_train(akmc_mlp_t m, const float *X, const float *T, int M, int epochs, int batch, float lr,
                   unsigned seed, double *loss_history);
/* descriptors of nq queries (atom = index into the batch's concatenated
 * atoms; vac, mover = 3 doubles each) at the batch's resident positions,
 * whose neighbour lists must be current (after akmc_batch_forces) */
int akmc_batch_descriptors(akmc_batch_t b, int nq, const int *atom, const double *vac, const double *mover,
                           float *X);

#ifdef __cplusplus
}
#endif
#endif
