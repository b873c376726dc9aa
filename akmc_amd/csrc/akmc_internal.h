// Internal device/host structures of libakmc_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AKMC_MAXSPEC 8
#define AKMC_MAXPAIR (AKMC_MAXSPEC * AKMC_MAXSPEC)
// neighbour entry: bits 0-19 local atom index, 20-23 / 24-27 / 28-31 the
// image shift (+7) along a, b, c.  Minimum-image lists store shift 0.
#define AKMC_NB_JMASK 0xFFFFFu
#define AKMC_NB_ZERO_SHIFT (0x777u << 20)
// spare 32-bit rows after a batch's neighbour lists (over-read by the
// unconditional entry prefetch of the 16-bit lists)
#define AKMC_NB_PAD_ROWS 64
#define AKMC_MAX_WALKER_ATOMS (1 << 20)
// workgroup size of the LDS-staged Sutton-Chen kernels (1024 / lpa atoms each)
#define AKMC_STAGE_BLOCK 1024
// active-zone bits of a search walker (BatchDev.zflag): zone on, a full
// FP32 / FP64 evaluation pending (the walker's first in that precision since
// its start), membership reset pending (set at the start)
#define AKMC_ZONE_ON 1
#define AKMC_ZONE_FULL32 2
#define AKMC_ZONE_FULL64 4
#define AKMC_ZONE_RESET 8
// cells per axis are capped so a walker's grid fits one workgroup's LDS
#define AKMC_MAX_CELLS_AXIS 16
#define AKMC_MAX_WALKER_CELLS (AKMC_MAX_CELLS_AXIS * AKMC_MAX_CELLS_AXIS * AKMC_MAX_CELLS_AXIS)

// Sutton-Chen pair parameters, FP32 for the fast path and FP64 for the
// verification path
template <typename T>
struct SCCoef {
    T a, n, m, eps;
    T rho_coef;  // (eps c)^2
    T fs_coef;   // n eps
    T fm_coef;   // 0.5 m (eps c)^2
};
struct SCPairDev {
    SCCoef<float> f;
    SCCoef<double> d;
    int n_int;   // n when n is a non-negative integer <= 64, else -1
    int m_int;
    int on;      // 1 when the species pair has a potential
};

// One tabulated EAM potential (MetalPotential, LAMMPS setfl layout read by
// readLEAM, potentials/metal/MetalPotential.cpp:276-361).  Every table is
// stored as (t_i, t_{i+1} - t_i) pairs so one 16-byte load feeds the linear
// interpolation t_i + ((t_{i+1} - t_i) (x - i h)) / h of eamMeshEnergy
// (potentials/metal/ManyBody.cpp:455-460) with the reference's operation
// order.  The derivative tables are the forward differences of
// calculateDerivativeMesh (MetalPotential.cpp:388-392), last entry 0.
#define AKMC_MAXEAM 8
struct EamTabDev {
    int nr, nrho;          // points of the r tables and of the embedding tables
    double dr, drho, cut;  // steps and the table's cutoff (potCutoff)
    const double2 *dens;   // [nr] rho(r)
    const double2 *pair;   // [nr][3] phi(r), phi'(r), rho'(r)
    const double2 *emb;    // [nrho][2] F(rho), F'(rho)
};

// short-range forms of the two-body table (TwoBody::shortrangeone,
// potentials/pair/TwoBody.h:173-276), parameters in the order of the
// potentials text: buck A rho C; morse D alpha r0 C12; ljones eps sigma;
// bhm A B sigma C D; severnsix eps sigma
enum : int { ION_BUCK = 0, ION_MORSE = 1, ION_LJONES = 2, ION_BHM = 3, ION_SEVNSIX = 4 };
struct IonPairDev {
    float A, inv_rho, C;
    double Ad, inv_rhod, Cd;
    int on;
    int form;
    float pf[6];
    double pd[6];
};

struct FieldDev {
    int kind;        // 0 many-body SC, 1 ionic
    int nspec;
    int min_image;
    int do_ewald;    // 0 none, 1 Ewald, 2 damped-shifted Coulomb
    double rc;
    double rc2;      // FP64 cutoff decision
    double rlist2;   // (rc + skin)^2
    double rzone2;   // (rc + skin/2 + 1e-6)^2: a free atom's list entries within it join its walker's zone
    double half_skin2;
    double rnear2;   // (r_near + skin)^2: list entries inside come first and
                     // are evaluated in FP64 on the FP32 path (0: none)
    double eta;
    double cfact;
    // do_ewald == 2: damped-shifted Coulomb on an erfc mesh (fieldErfcGen,
    // potentials/pair/Field.cpp:19-73): points, spacing, shift terms
    const double *dsf_e, *dsf_f;
    int dsf_n;
    double dsf_del, dsf_rfld0, dsf_rfld1;
    SCPairDev sc[AKMC_MAXPAIR];
    IonPairDev ion[AKMC_MAXPAIR];
    // tabulated EAM (kind 0 with eam = 1): species pair -> table (-1 none)
    int eam;
    int eam_one;           // one table serves every species pair
    int eam_tab[AKMC_MAXPAIR];
    EamTabDev eamt[AKMC_MAXEAM];
};
// bits of BatchDev::overflow
#define AKMC_OV_LIST 1     // a neighbour list ran out of slots
#define AKMC_OV_EAM_RHO 2  // a density fell outside its embedding table

struct WalkerDev {
    int off, n;          // first atom, atom count
    int cell_off;        // first cell of this walker's grid
    int nc[3];           // cells per axis
    int shells[3];       // explicit image shells (no-image fields)
    int k_off, nk;       // slice of the k-vector arrays
    int gmax;            // largest |n| along any axis of this walker's k set
    double L[3], invL[3];
};

// per-walker outputs of one force evaluation (FP64)
struct WalkerOut {
    double total, rcp, real, many, vdw, self;
    double stress[6];
};

struct BatchDev {
    int A, W, maxnb, ncells_total, nk_total, G1, maxnk;
    int lpa;              // lanes per atom of the Sutton-Chen pair kernels (1, 2, 4)
    int nb16;             // list entries are 16-bit local indices (minimum-image
                          // many-body lists of walkers below 65536 atoms)
    int unroll;           // neighbour gathers in flight per lane (1, 2, 4, 8)
    int staged;           // Sutton-Chen pair kernels read the gather records of
                          // a walker from LDS (one workgroup per chunk of atoms)
    const int2 *chunk_tab[4];  // staged chunk -> (walker, first local atom) at
    int nchunk_tab[4];         // 1, 2, 4, 8 lanes per atom (1024 / lanes atoms each)
    int stage_bytes;           // dynamic LDS of a staged launch
    const FieldDev *field;
    const WalkerDev *walkers;
    const int *wk;        // atom -> walker
    double *x, *y, *z;    // positions (FP64 master copy)
    double4 *p4;          // packed (x, y, z, type) of every atom for the gathers
    double *xr, *yr, *zr; // positions at the last list build
    const int *type;
    const double *q;
    double *fx, *fy, *fz;
    uint32_t *nbr;        // [maxnb][A] column-major
    int *nnb;
    int *nnear;           // leading list entries inside the near shell
    int *rebuild;         // per walker flag
    int *overflow;        // sticky capacity flag
    int *cell_count, *cell_start, *cell_atoms;
    const int *cell_wk;   // cell -> walker
    int *cell_work;       // cells whose lists are rebuilt this step (queued by k_cell_bin)
    int *cell_work_n;
    double *rho;          // SC density
    float *inv_sqrt_rho;  // SC 1/sqrt(rho), -1 when rho < 1e-8 (FP32 walkers)
    double *isr64;        // the same in FP64 (FP64 walkers)
    double *e_a, *e_b;    // per-atom energies (SC: many, pair; ionic: real, vdw)
    double *e_c;          // ionic: reciprocal per atom
    double *st_atom;      // [6][A] per-atom stress when requested
    // Ewald
    const int *kn;        // [3][nk_total] integer k indices
    const double *kv;     // [3][nk_total] k vectors
    const double *g0, *g1;
    double2 *ptab;        // [A][3][G1] e^{i n theta}
    double *Ck, *Sk;      // structure factors
    WalkerOut *out;
    int *active;          // per walker: 0 skip, 1 FP32 pair terms, 2 FP64
    int energy_only;      // this launch evaluates energies only (Field::calculateEnergy):
                          // no force pass, the energy-only short-range forms
    unsigned long long *rebuild_count;  // list rebuilds, counted by the list kernels
    // per-atom energies of the FP64 Sutton-Chen kernels (null: e_a, e_b), so
    // each precision keeps its own values of the atoms it did not visit
    double *e_a64, *e_b64;
    // Active zone of the saddle-search walkers of a Sutton-Chen field (null
    // outside search micro-steps).  A search moves only its free atoms, so
    // a force evaluation needs the forces on the free atoms, which need the
    // densities of the atoms within the cutoff of a free atom, and an energy
    // change involves only those atoms' energies.  A zoned walker's pass 2
    // visits its free atoms, its pass 1 the zone: the free atoms and every
    // atom on a free atom's Verlet list (rc + skin, so the zone holds every
    // atom within rc of a free atom until the next rebuild), kept as a union
    // over the rebuilds since the walker started.  Every other atom has had
    // no free atom within rc since the walker's first full evaluation in
    // each precision, so its energy, kept from that evaluation, is the one a
    // full evaluation would give, bit for bit.
    // search micro-steps on minimum-image Sutton-Chen fields: pass 1 also
    // writes each atom's pair energy and pass 2 writes none, so the pass-2
    // variant may differ from pass 1's without changing any energy
    int pe1;
    int lpa2, unroll2, staged2;  // pass-2 variant (0: the batch's)
    int lpa64, unroll64, staged64;  // FP64 variant of mixed steps (lpa64 0: 4 x 4, unstaged)
    const int *zfrozen;   // [A] the search's frozen mask (zone marking in the list build)
    int *zflag;           // [W] AKMC_ZONE_* bits
    int *zmark;           // [A] zone membership
    int *zlist;           // [A] walker w's zone (local ids, ascending) at zlist[off_w ..)
    int *nz;              // [W] zone size
    const int *zact;      // [A] walker w's free atoms (global ids) at zact[off_w ..)
    const int *znact;     // [W] free atoms per walker
    // work lists of the search micro-steps' Sutton-Chen pair kernels
    // (k_work_list): FP32 pass 1, FP32 pass 2, FP64 pass 1, FP64 pass 2
    int2 *wl_tab[4];      // (walker, first slot) per chunk
    int *wl_count;        // [4] chunks per list
    int wl_grid_staged;   // persistent grids of those kernels (multiples of 8)
    int wl_grid_plain;
};

// entry k of atom i's neighbour list in the 32-bit form (local index and
// image shift), whichever width the batch stores
__device__ inline uint32_t nbr_entry(const BatchDev &b, int k, int i)
{
    // batches keep A below 2^24 and (maxnb + pad) A below 2^32
    const uint32_t o = __umul24(uint32_t(k), uint32_t(b.A)) + uint32_t(i);
    return b.nb16 ? (uint32_t(reinterpret_cast<const uint16_t *>(b.nbr)[o]) | AKMC_NB_ZERO_SHIFT) : b.nbr[o];
}

// launcher (field_kernels.hip): list maintenance + one force evaluation of
// every walker; hf is the host copy of *b.field
void akmc_launch_forces_impl(const BatchDev &b, const FieldDev &hf, int fp64, int want_stress,
                             hipStream_t s);
void akmc_launch_phase(const BatchDev &b, const FieldDev &hf, int fp64, int want_stress, int phase,
                       hipStream_t s);
extern "C" unsigned long long *akmc_rebuild_counter_ptr();

// what the allocation-free force entry needs of a field handle (capi.cpp)
struct FieldView {
    const FieldDev *dev;   // device copy of the field tables (set up by akmc_field_setup)
    const FieldDev *host;  // host copy (launch decisions)
    double rc, skin;
    int kind, min_image, eam, fp64;
};
int akmc_field_view(struct akmc_field_s *f, FieldView *v);
// records the message akmc_last_error() returns and passes code through
extern "C" int akmc_fail(int code, const char *fmt, ...);

// ------------------------------------------------------------ searches

// walker kinds and states of the device state machines
// WK_BASIN2: FIRE relax of a kicked converged saddle into the basin beyond it
enum : int { WK_IDLE = 0, WK_RELAX = 1, WK_DIMER = 2, WK_VERIFY = 3, WK_ART = 4, WK_BASIN2 = 5 };
enum : int {
    ST_DONE = 0,
    ST_R_INIT = 1,   // relax: force at X (before the loop)
    ST_R_STEP = 2,   // relax: force at the updated X
    ST_R_E64 = 3,    // relax: FP64 energy at the relaxed X
    ST_D_F0 = 10,    // dimer: force at the centre
    ST_D_F1 = 11,    // dimer: force at the end X + D n
    ST_D_F1T = 12,   // dimer: force at the trial end X + D n'
    ST_V_F0 = 20,    // verify: FP64 force at the saddle
    ST_V_LZ = 21,    // verify: FP64 force at X + delta q_k (Lanczos)
    ST_V_LZM = 22,   // verify: FP64 force at X - delta q_k
    ST_A_L0 = 30,    // ART: force at X, start of a Lanczos evaluation
    ST_A_LZ = 31,    // ART: force at X + deltaX q_i
};
#define AKMC_LZ_MAX 32

// FIRE relax parameters (RelaxCntrl, Common/RelaxCntrl.h:15-48)
struct RelaxParamsDev {
    double normTol, forceTol, initTimeStep, maxTimeStep, alphaStart, alphaDecrease, stepIncrease,
        stepDecrease, timeStepUnit, minTimeStep;
    int maxIter, N_min;
    int method;   // 0 FIRE, 1 FIRE2, 2 force and energy only (AKMC_RELAX_*)
    int verlet;   // velocity-Verlet instead of semi-implicit Euler
    int N_delaySteps, N_maxUphill, initialDelay;
};

// ART parameters (ARTCntrl names, Common/ArtCntrl.h:14-74)
struct ARTParamsDev {
    double deltaX, maxEigenvalue, eigValTol, parDamp, normTol, forceTol, initTimeStep, alphaDecrease,
        stepIncrease, stepDecrease, alphaStart, maxTimeStep, timeStepUnit;
    int numLanczosVectors, maxParIter, N_min;
};

// dimer parameters (DimerCntrl names, Common/DimerCntrl.h:18-88)
struct DimerParamsDev {
    double deltaTheta, dimerDisplacement, maxEigenvalue, normTol, forceTol, alphaStart, alphaDecrease,
        stepIncrease, stepDecrease, initTimeStep, maxTimeStep, maxStep, timeStepUnit;
    int maxRotations, maxParIter, N_min;
};

struct WalkerState {
    int kind, state, it, nrot, status, nstep, fcalls, lk;
    double E0, C, g, t1, dt, alpha, E_last;
    // FP64 verification of a converged dimer: search energy/curvature,
    // FP64 energy, max FP64 force on free atoms, two lowest Ritz values
    double e32, c32, e64, fmax64, lam1, lam2;
    int verified_state;  // 0 none, 1 verified data present
    int inloop;          // ART: 0 = the Lanczos before the FIRE loop
    // ART: norms of the modified forces before the move, F.v, half step,
    // last and previous Lanczos eigenvalue, Lanczos beta
    double fnpar, fnperp, fmpar, fmperp, dotfv, half, lam, lamold, lbeta;
    int lconv;
    int nuphill;  // FIRE2: uphill steps in a row
    // relax of a walker with frozen atoms: their common velocity (they all
    // start at rest and get the same centre-of-mass corrections and resets)
    double vf[3];
    // basin 2 of a converged saddle: energy (search precision, FP64),
    // largest squared displacement from the basin, new-basin flag, FIRE
    // status, force calls and iterations of its relax
    double eb2, eb2_64, b2_maxd2;
    int b2_new, b2_status, b2_calls, b2_iters;
};

struct SearchDev {
    double *X, *N, *V, *F0, *F1, *T;  // [3][A] each
    const int *frozen;                // [A]
    const int *act;                   // [A]: walker w's free atoms (global ids) at act[off_w ..)
    const int *nact;                  // [W]: number of free atoms per walker
    WalkerState *ws;                  // [W]
    unsigned long long *evals32;      // [W]: FP32 force evaluations consumed by the state machines
    RelaxParamsDev rp;
    DimerParamsDev dp;
    ARTParamsDev ap;
    // Lanczos basis Q [LZ_MAX][3][A]; per walker alpha, beta and the Ritz
    // vector: lz [W][3*LZ_MAX]
    double *Q;
    double *lz;
    int verify_nl;
    double verify_delta;
    int default_mode;                 // active mode of searching walkers (1 FP32, 2 FP64)
    // basin 2 after every converged search (checkForNewBasin,
    // kmc/KineticMonteCarlo.cpp:1626-1713): on when basin2_on and basin set
    const double *basin;              // [3][A] the walkers' basin positions
    double *Xs;                       // [3][A] saddle positions kept for the host
    double basin2_delta, basin_r2, max_r2;
    int basin2_on;
    int final64;                      // relaxes end with one FP64 energy evaluation
};

void akmc_launch_search_logic(const BatchDev &b, const SearchDev &s, hipStream_t st);
// the same state machines in 1024-thread workgroups (large walkers)
void akmc_launch_search_logic_wide(const BatchDev &b, const SearchDev &s, hipStream_t st);
// one walker start packed for akmc_launch_load_walkers: doubles at dbl
// (mode [3][n], then positions [3][n] when has_x), ints at ints (frozen [n],
// then the free-atom list [nact])
struct LoadItem {
    int walker, has_x, nact, pad;
    long long dbl, ints;
};
// the listed walkers' slices of three [A] arrays packed into out (device)
void akmc_launch_gather_walkers(const WalkerDev *walkers, const double *const src[3], const int *ws,
                                const long long *dst, int nw, double *out, hipStream_t st);
void akmc_launch_load_walkers(const BatchDev &b, const SearchDev &s, const LoadItem *items, int nitems,
                              const double *D, const int *I, hipStream_t st);
void akmc_launch_probe_init(const BatchDev &b, const SearchDev &s, hipStream_t st);
// a walker start: its fresh state, zone flags and active flag
struct StartItem {
    int walker, zflag, active, pad;
    WalkerState ws;
};
// scatters the starts and asks each started walker for a list rebuild on
// its first micro-step (zflag and harvested may be null)
void akmc_launch_start_walkers(const BatchDev &b, const SearchDev &s, const StartItem *items, int nitems, int *zflag,
                               int *harvested, hipStream_t st);
// copies the positions X of every walker that has ended and is not yet
// harvested into out ([3][A], host memory mapped for the device) and marks it
void akmc_launch_harvest_pack(const BatchDev &b, const SearchDev &s, int *harvested, double *out, hipStream_t st);

// learned proposal (mlp_kernels.hip)
void akmc_mlp_launch_weights(const float *p, uint16_t *w1t, uint16_t *w2t, uint16_t *w3t, uint16_t *w2r,
                             uint16_t *w3r, hipStream_t s);
void akmc_mlp_launch_forward(const float *X, int M, const float *p, const uint16_t *w1t, const uint16_t *w2t,
                             const uint16_t *w3t, float *Z1, float *H1, float *Z2, float *H2, float *Y, hipStream_t s);
void akmc_mlp_launch_backward(const float *X, const float *T, int M, const uint16_t *w2r, const uint16_t *w3r,
                              const float *Z1, const float *H1, const float *Z2, const float *H2, const float *Y,
                              float *dY, float *dZ2, float *dZ1, double *loss_part, float *P, int rows_per_chunk,
                              float *grad, hipStream_t s);
void akmc_mlp_launch_adam(float *p, const float *g, float *m1, float *m2, int n, float lr, int t, hipStream_t s);
void akmc_mlp_launch_gather(const float *src, int width, const int *idx, int M, float *dst, hipStream_t s);
void akmc_mlp_launch_loss_accum(const double *part, int nb, double scale, double *acc, hipStream_t s);
void akmc_mlp_launch_descriptors(const BatchDev &b, int nq, const int *qatom, const double *qvac,
                                 const double *qmov, float *X, hipStream_t s);
