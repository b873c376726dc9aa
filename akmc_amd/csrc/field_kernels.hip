// Force-field kernels for gfx950: neighbour lists, Sutton-Chen two-pass
// EAM, Buckingham + Ewald real space, Ewald reciprocal space.
//
// Data layout (all walkers of a batch concatenated, SoA, FP64 positions):
//   x/y/z[A], type[A], q[A]; neighbour list nbr[maxnb][A] (column-major, so
//   lane i of a wavefront reads entry k of atoms i..i+63 in one coalesced
//   access); per-walker grids of cells for the minimum-image list build.
// Precision: the cutoff test r^2 <= rc^2 and the minimum-image displacement
// are FP64; pair physics is T (float on the fast path, double for FP64
// verification); per-atom force/energy sums are FP64.
#include "akmc_internal.h"

#include <math.h>

namespace {

constexpr int kBlock = 256;

__device__ inline int walker_of(const BatchDev &b, int i) { return b.wk[i]; }

// Workgroups are dealt round-robin to the 8 XCDs; remapping the hardware
// workgroup id so that each XCD owns one contiguous range of atoms keeps a
// walker's positions in a single XCD's L2.  Launches use xcd_grid(n).
constexpr int kXCD = 8;
__host__ __device__ inline int xcd_grid(int nblocks) { return ((nblocks + kXCD - 1) / kXCD) * kXCD; }
__device__ inline int xcd_atom(int nblocks)
{
    const int per = (nblocks + kXCD - 1) / kXCD;
    const int logical = (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
    return logical < nblocks ? logical * blockDim.x + threadIdx.x : -1;
}
__device__ inline bool skipped(const BatchDev &b, int w) { return b.active[w] == 0; }
// the pair-physics kernels of type T serve the walkers whose precision mode
// matches (1 FP32, 2 FP64)
template <typename T>
__device__ inline bool other_precision(const BatchDev &b, int w)
{
    return b.active[w] != (sizeof(T) == 8 ? 2 : 1);
}

// the per-atom energy arrays the many-body kernels of type T write
template <typename T>
__device__ inline double *energy_a(const BatchDev &b)
{
    return (sizeof(T) == 8 && b.e_a64) ? b.e_a64 : b.e_a;
}
template <typename T>
__device__ inline double *energy_b(const BatchDev &b)
{
    return (sizeof(T) == 8 && b.e_b64) ? b.e_b64 : b.e_b;
}

// walker w evaluates only its active zone in a pass of type T: a search
// micro-step, the zone on, and the walker's full evaluation in this
// precision done
template <typename T>
__device__ inline bool zoned(const BatchDev &b, int w)
{
    if (!b.zflag) return false;
    const int f = b.zflag[w];
    return (f & AKMC_ZONE_ON) && !(f & (sizeof(T) == 8 ? AKMC_ZONE_FULL64 : AKMC_ZONE_FULL32));
}

template <typename T>
__device__ inline const SCCoef<T> &sc_coef(const SCPairDev &p);
template <>
__device__ inline const SCCoef<float> &sc_coef<float>(const SCPairDev &p) { return p.f; }
template <>
__device__ inline const SCCoef<double> &sc_coef<double>(const SCPairDev &p) { return p.d; }

// 1/x: the hardware reciprocal (1 ulp) on the FP32 path, an exact division
// on the FP64 verification path
template <typename T>
__device__ inline T recip(T x) { return T(1) / x; }
template <>
__device__ inline float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// u^x for non-integer exponents: libm pow on the FP64 path; on the FP32
// path the integer part of x by squaring and the fraction by libm log2f /
// exp2f (relative-accurate near u = 1, where (a/r)^2 lives), which keeps
// the error of u^x at a few ulp without the cost of a full powf
__device__ inline double pow_general(double u, double x) { return pow(u, x); }
__device__ inline float pow_general(float u, float x)
{
    if (!(x >= 0.0f && x < 64.0f)) return exp2f(x * log2f(u));  // outside the split's range
    const float xi = floorf(x);
    float r = 1.0f, base = u;
    for (int p = int(xi); p > 0; p >>= 1) {
        if (p & 1) r *= base;
        base *= base;
    }
    return r * exp2f((x - xi) * log2f(u));
}

// u2^(k/2) for an integer k >= 0 by repeated squaring; with k uniform across
// the wavefront the loop is scalar control flow around the multiplications
template <typename T>
__device__ inline T pow_int(T u2, int k)
{
    T r = T(1), base = u2;
    int p = k >> 1;
    while (p) {
        if (p & 1) r *= base;
        base *= base;
        p >>= 1;
    }
    if (k & 1) r *= sqrt(u2);
    return r;
}

// square root for the integer powers: the hardware instruction (1 ulp) on
// the FP32 path, whose arguments (a/r)^2 are normal numbers far from the
// denormal range the correctly rounded expansion guards; libm on FP64
__device__ inline double root(double x) { return sqrt(x); }
__device__ inline float root(float x) { return __builtin_amdgcn_sqrtf(x); }

// 1/sqrt(x) in FP64 for normal x > 0: the hardware estimate refined by two
// Newton steps y += y (1 - x y^2) / 2 (quadratic convergence), a few ulp
// from the correctly rounded value at a third of the cost of a division and
// a square root
__device__ inline double rsqrt_fp64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double r = fma(-(x * y), y, 1.0);
    y = fma(0.5 * y, r, y);
    r = fma(-(x * y), y, 1.0);
    return fma(0.5 * y, r, y);
}

// The single-species power inputs of U pairs: ir2 = 1/r^2, u2 = (a/r)^2 and,
// in FP64, rt = a/r.  FP32: hardware reciprocal (the power takes sqrt(u2)
// from the hardware square root for odd exponents).  FP64: one refined 1/r,
// so that neither a division nor a square root is needed.
template <typename T, int U>
__device__ inline void sc_powers_in(const double (&r2)[U], T a, T (&ir2)[U], T (&u2)[U], T (&rt)[U])
{
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double ir = rsqrt_fp64(r2[u]);
            ir2[u] = ir * ir;
            u2[u] = a * a * ir2[u];
            rt[u] = a * ir;
        }
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ir2[u] = recip<T>(T(r2[u]));
            u2[u] = a * a * ir2[u];
            rt[u] = T(0);  // the FP32 power takes its root from u2 itself
        }
    }
}

// (a/r)^9 and (a/r)^6 of the single-pair kernels from u2 = (a/r)^2, with
// the compile-time exponents of the Sutton-Chen pairs they serve: the
// products of the repeated squaring, u8 sqrt(u2) and u2 u4, without its
// loop
template <typename T, int U>
__device__ inline void pow_9_6(const T (&u2)[U], const T (&rt)[U], T (&u9)[U], T (&u6)[U])
{
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const T u4 = u2[u] * u2[u];
        const T u8 = u4 * u4;
        u9[u] = u8 * (sizeof(T) == 8 ? rt[u] : root(u2[u]));
        u6[u] = u2[u] * u4;
    }
}
template <typename T, int U>
__device__ inline void pow_6(const T (&u2)[U], T (&u6)[U])
{
#pragma unroll
    for (int u = 0; u < U; ++u) u6[u] = u2[u] * (u2[u] * u2[u]);
}

template <typename T>
__device__ inline T pow_half(T u2, int k, T e)
{
    // u2^(e/2) with u2 = (a/r)^2; exact repeated squaring when e is an integer
    if (k >= 0) {
        T r = T(1), base = u2;
        int p = k >> 1;
        while (p) {
            if (p & 1) r *= base;
            base *= base;
            p >>= 1;
        }
        if (k & 1) r *= sqrt(u2);
        return r;
    }
    return pow_general(u2, e * T(0.5));
}

__device__ inline void min_image(double &dx, double &dy, double &dz, const WalkerDev &w)
{
    dx -= w.L[0] * rint(dx * w.invL[0]);
    dy -= w.L[1] * rint(dy * w.invL[1]);
    dz -= w.L[2] * rint(dz * w.invL[2]);
}

__device__ inline void entry_disp(uint32_t e, double &dx, double &dy, double &dz, const WalkerDev &w,
                                  bool minimg)
{
    if (minimg) {
        min_image(dx, dy, dz, w);
    } else {
        int sa = int((e >> 20) & 15u) - 7, sb = int((e >> 24) & 15u) - 7, sc = int((e >> 28) & 15u) - 7;
        dx -= sa * w.L[0];
        dy -= sb * w.L[1];
        dz -= sc * w.L[2];
    }
}

// ---------------------------------------------------------------- lists

// Per atom of an active walker, in one pass over the positions: the gather
// copy of the positions and species (one 32-byte record per neighbour
// read), and the skin test that flags the walker's list for a rebuild when
// the atom moved more than skin/2 since the list was built.
__global__ void k_pack_check(BatchDev b)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *b.cell_work_n = 0;  // k_cell_bin queues this step's cells
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.A) return;
    int w = walker_of(b, i);
    if (skipped(b, w)) return;
    // an atom outside a running search's zone has not moved since the
    // walker's start and its record is only read by full evaluations, whose
    // pass 1 rewrites the fourth word it reads (single pair type)
    if (b.zflag && (b.zflag[w] & (AKMC_ZONE_ON | AKMC_ZONE_RESET)) == AKMC_ZONE_ON && !b.zmark[i]) return;
    const double x = b.x[i], y = b.y[i], z = b.z[i];
    b.p4[i] = make_double4(x, y, z, (double)b.type[i]);
    double dx = x - b.xr[i], dy = y - b.yr[i], dz = z - b.zr[i];
    if (dx * dx + dy * dy + dz * dz > b.field->half_skin2) b.rebuild[w] = 1;
    // a walker's fresh zone is marked by its list build (its start already
    // asks for one)
    if (b.zflag && (b.zflag[w] & AKMC_ZONE_RESET)) b.rebuild[w] = 1;
}

// k_pack_check for batches with zones (search micro-steps): one workgroup
// per walker, over its zone when the zone is established (the other atoms
// have not moved since the walker's start), else over all its atoms
constexpr int kPackBlock = 1024;
__global__ void __launch_bounds__(kPackBlock) k_pack_check_zoned(BatchDev b)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *b.cell_work_n = 0;  // k_cell_bin queues this step's cells
    const int w = blockIdx.x;
    if (skipped(b, w)) return;
    const WalkerDev W = b.walkers[w];
    const int f = b.zflag[w];
    const bool zoned = (f & (AKMC_ZONE_ON | AKMC_ZONE_RESET)) == AKMC_ZONE_ON;
    const int n = zoned ? b.nz[w] : W.n;
    const double h2 = b.field->half_skin2;
    bool moved = false;
    for (int k = threadIdx.x; k < n; k += kPackBlock) {
        const int i = W.off + (zoned ? b.zlist[W.off + k] : k);
        const double x = b.x[i], y = b.y[i], z = b.z[i];
        b.p4[i] = make_double4(x, y, z, (double)b.type[i]);
        const double dx = x - b.xr[i], dy = y - b.yr[i], dz = z - b.zr[i];
        moved |= dx * dx + dy * dy + dz * dz > h2;
    }
    // a walker's fresh zone is marked by its list build (its start already
    // asks for one)
    if (moved || (f & AKMC_ZONE_RESET)) b.rebuild[w] = 1;
}

// record the positions the lists are built from (positions are never
// wrapped on the device: the minimum-image and explicit-image kernels are
// periodic in them)
__global__ void k_ref_positions(BatchDev b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.A) return;
    int w = walker_of(b, i);
    if (!b.rebuild[w]) return;
    b.xr[i] = b.x[i];
    b.yr[i] = b.y[i];
    b.zr[i] = b.z[i];
}

__device__ inline int cell_coord(double x, double invL, int nc)
{
    double s = x * invL;
    s -= floor(s);
    int c = int(s * nc);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// walker-local cell of atom i
__device__ inline int atom_cell_of(const BatchDev &b, const WalkerDev &W, int i)
{
    const int ca = cell_coord(b.x[i], W.invL[0], W.nc[0]);
    const int cb = cell_coord(b.y[i], W.invL[1], W.nc[1]);
    const int cc = cell_coord(b.z[i], W.invL[2], W.nc[2]);
    return (ca * W.nc[1] + cb) * W.nc[2] + cc;
}

// k_cell_bin: one workgroup of kBinBlock threads per rebuilt walker; the
// cells of walkers up to kBinAtoms atoms stay in LDS between the counting
// and the placement
constexpr int kBinBlock = 1024;
constexpr int kBinAtoms = 8192;

// home atoms per list-build work item (k_nbr_build_cells)
constexpr int kNbSlotsQ = 16;

// Cell grid of one walker per workgroup (walkers whose list is rebuilt):
// counts, exclusive scan and a stable placement of the atoms in index order,
// so every cell lists its atoms by ascending index and the neighbour lists
// (and the FP64 sums that follow them) do not depend on scheduling.
// inclusive scan of v over the kBinBlock threads: wavefront shuffles, then
// the wavefront totals in order (two barriers); total gets the sum
__device__ inline int bin_scan(int v, int *wsum, int &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(x, o);
        if (lane >= o) x += t;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int before = 0, all = 0;
    for (int k = 0; k < kBinBlock / 64; ++k) {
        const int t = wsum[k];
        if (k < wave) before += t;
        all += t;
    }
    __syncthreads();  // the next scan rewrites wsum
    total = all;
    return x + before;
}

__global__ void __launch_bounds__(kBinBlock) k_cell_bin(BatchDev b)
{
    __shared__ int fill[AKMC_MAX_WALKER_CELLS];
    __shared__ int chunk[kBinBlock], rank_[kBinBlock], gsize[kBinBlock];
    __shared__ int acell[kBinAtoms];  // each atom's cell, kept from the counts (walkers up to kBinAtoms)
    __shared__ int wsum[kBinBlock / 64];
    const int w = blockIdx.x;
    if (!b.rebuild[w]) return;
    const WalkerDev W = b.walkers[w];
    const int nc = W.nc[0] * W.nc[1] * W.nc[2];
    const int tid = threadIdx.x;
    for (int c = tid; c < nc; c += kBinBlock) fill[c] = 0;
    // a restarted walker's zone starts empty (the list build marks it)
    if (b.zflag && (b.zflag[w] & AKMC_ZONE_RESET))
        for (int l = tid; l < W.n; l += kBinBlock) b.zmark[W.off + l] = 0;
    __syncthreads();
    const bool keep = W.n <= kBinAtoms;
    // the counts, and the positions the walker's lists are built from
    for (int l = tid; l < W.n; l += kBinBlock) {
        const int i = W.off + l;
        const double x = b.x[i], y = b.y[i], z = b.z[i];
        b.xr[i] = x;
        b.yr[i] = y;
        b.zr[i] = z;
        const int cell = (cell_coord(x, W.invL[0], W.nc[0]) * W.nc[1] + cell_coord(y, W.invL[1], W.nc[1])) * W.nc[2] +
                         cell_coord(z, W.invL[2], W.nc[2]);
        if (keep) acell[l] = cell;
        atomicAdd(&fill[cell], 1);
    }
    __syncthreads();
    // exclusive scan of the counts; fill[c] becomes the cell's next free slot
    int carry = W.off;
    for (int base = 0; base < nc; base += kBinBlock) {
        const int c = base + tid;
        const int v = c < nc ? fill[c] : 0;
        int tot;
        const int incl = bin_scan(v, wsum, tot);
        if (c < nc) {
            const int start = carry + incl - v;
            b.cell_count[W.cell_off + c] = v;
            b.cell_start[W.cell_off + c] = start;
            fill[c] = start;
        }
        carry += tot;
    }
    __syncthreads();
    // stable placement in index order, kBinBlock atoms at a time: each
    // wavefront finds, by ballots, the lanes of its 64 atoms that hold the
    // same cell (rank below, group size, highest lane); then one wavefront
    // walks the groups of 64 in order: an atom's slot is the cell's fill
    // plus its rank, and the highest lane of each group advances the fill
    for (int base = 0; base < W.n; base += kBinBlock) {
        const int l = base + tid;
        const int c = l < W.n ? (keep ? acell[l] : atom_cell_of(b, W, W.off + l)) : -1;
        unsigned long long todo = __ballot(c >= 0);
        unsigned long long same = 0;
        while (todo) {
            const int leader = __ffsll((long long)todo) - 1;
            const int cl = __shfl(c, leader);
            const unsigned long long grp = __ballot(c == cl);
            if (c == cl) same = grp;
            todo &= ~grp;
        }
        const int wl = tid & 63;
        chunk[tid] = c;
        rank_[tid] = __popcll(same & ((1ull << wl) - 1ull));
        gsize[tid] = (c >= 0 && (same >> wl) == 1ull) ? __popcll(same) : 0;  // set on the group's highest lane
        __syncthreads();
        if (tid < 64) {
            for (int wv = 0; wv < kBinBlock / 64; ++wv) {
                const int k = wv * 64 + tid;
                const int ck = chunk[k];
                if (ck >= 0) {
                    b.cell_atoms[fill[ck] + rank_[k]] = W.off + base + k;
                    if (gsize[k]) fill[ck] += gsize[k];
                }
            }
        }
        __syncthreads();
    }
    // queue this walker's list-build items: home chunks of kNbSlots atoms per
    // cell (chunk indices above 255 are served by the item of chunk % 256),
    // in cell order.  After the placement fill[c] is the end of cell c, so a
    // cell's count is fill[c] minus the end of cell c - 1; the items per
    // cell are scanned across the workgroup and the walker takes its queue
    // range with one atomic.
    __shared__ int qbase;
    auto items_of = [&](int c) {
        const int n = fill[c] - (c == 0 ? W.off : fill[c - 1]);
        const int k = (n + kNbSlotsQ - 1) / kNbSlotsQ;
        return k < 256 ? k : 256;
    };
    int mine = 0;
    for (int c = tid; c < nc; c += kBinBlock) mine += items_of(c);
    int items;
    (void)bin_scan(mine, wsum, items);
    if (tid == 0) qbase = atomicAdd(b.cell_work_n, items);
    __syncthreads();
    carry = 0;
    for (int base = 0; base < nc; base += kBinBlock) {
        const int c = base + tid;
        const int v = c < nc ? items_of(c) : 0;
        int tot;
        const int incl = bin_scan(v, wsum, tot);
        if (c < nc) {
            const int at = qbase + carry + incl - v;
            for (int ch = 0; ch < v; ++ch) b.cell_work[at + ch] = ((W.cell_off + c) << 8) | ch;
        }
        carry += tot;
    }
}

// The entry after atom i's list (row n <= maxnb; the buffer holds
// AKMC_NB_PAD_ROWS spare rows) is i's own "no neighbour" entry: local index
// i, zero image shift.  The pair kernels' entry prefetch clamps its row to
// the list length and reads this entry for every slot past the end, whose
// separation 0 the r^2 floor rejects.
__device__ inline void put_sentinel(const BatchDev &b, int i, int n, int off)
{
    const size_t o = (size_t)n * b.A + i;
    const uint32_t e = uint32_t(i - off) | AKMC_NB_ZERO_SHIFT;
    if (b.nb16) reinterpret_cast<uint16_t *>(b.nbr)[o] = uint16_t(e & 0xFFFFu);
    else b.nbr[o] = e;
}

__device__ inline void push_nbr(const BatchDev &b, int i, int &cnt, uint32_t e)
{
    if (cnt < b.maxnb) {
        const size_t o = (size_t)cnt * b.A + i;
        if (b.nb16) reinterpret_cast<uint16_t *>(b.nbr)[o] = uint16_t(e & 0xFFFFu);
        else b.nbr[o] = e;
    } else {
        *b.overflow = 1;
    }
    ++cnt;
}

// Minimum-image Verlet list.  A work item is up to kNbSlots home atoms of
// one cell of a walker whose list is rebuilt (queued by k_cell_bin); a
// workgroup stages the item's candidates -- the atoms of the neighbouring
// cells (c-1..c+1 per axis, every cell once along axes with fewer than 3
// cells), in cell order and ascending index within a cell -- into LDS in
// one cooperative pass (every thread's loads in flight together), up to
// kNbCap at a time.  kNbLanes lanes serve one home atom: they test
// kNbLanes * kNbIlp consecutive candidates per round and place their hits
// by counts in the wavefront's hit ballot, so entries come out exactly in candidate
// order; with a near shell (ionic fields) all entries inside it come first.
// Entries are every j != i with min-image distance <= rc + skin.
constexpr int kNbLanes = 16;
constexpr int kNbSlots = kBlock / kNbLanes;  // home atoms per work item
static_assert(kNbSlots == kNbSlotsQ, "queued items and the list build agree on the chunk size");
constexpr int kNbIlp = 4;                    // candidates per lane per round
constexpr int kNbCap = 1024;                 // candidates staged in LDS at a time (28 KB)
constexpr int kNbStage = kNbCap / kBlock;    // candidates per thread in one staging
constexpr int kNbWaves = kBlock / 64;
static_assert(kNbStage * kBlock == kNbCap, "the staging covers kNbCap candidates");

__device__ inline void put_nbr(const BatchDev &b, int i, int k, uint32_t e)
{
    if (k < b.maxnb) {
        const size_t o = (size_t)k * b.A + i;
        if (b.nb16) reinterpret_cast<uint16_t *>(b.nbr)[o] = uint16_t(e & 0xFFFFu);
        else b.nbr[o] = e;
    } else {
        *b.overflow = 1;
    }
}

// one work item: home atoms [chunk * kNbSlots, +kNbSlots) of cell cg
__device__ void nbr_build_item(const BatchDev &b, int cg, int chunk, double *sx, double *sy, double *sz, int *sj,
                               int *c_start, int *c_pref, int &n_cells, double *box, int *scnt)
{
    const int w = b.cell_wk[cg];
    const WalkerDev W = b.walkers[w];
    // the home atom's loads go out before the candidate cells are set up
    const int s0 = b.cell_start[cg], n0 = b.cell_count[cg];
    const int q = chunk * kNbSlots + threadIdx.x / kNbLanes;
    const bool have = q < n0;  // slot 0 always has an atom
    const int i = have ? b.cell_atoms[s0 + q] : 0;
    const double xi = have ? b.x[i] : 0.0, yi = have ? b.y[i] : 0.0, zi = have ? b.z[i] : 0.0;
    // candidate cells, one per lane of the first wavefront in (a, b, c)
    // order, with an inclusive scan of their counts
    if (threadIdx.x < 64) {
        const int t = threadIdx.x;
        const int lc = cg - W.cell_off;
        const int cen[3] = {lc / (W.nc[1] * W.nc[2]), (lc / W.nc[2]) % W.nc[1], lc % W.nc[2]};
        int lo[3], span[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = W.nc[a] >= 3 ? cen[a] - 1 : 0;
            span[a] = W.nc[a] >= 3 ? 3 : W.nc[a];
        }
        const int ncell = span[0] * span[1] * span[2];
        int st = 0, v = 0;
        if (t < ncell) {
            const int pa = lo[0] + t / (span[1] * span[2]), pb = lo[1] + (t / span[2]) % span[1];
            const int pc = lo[2] + t % span[2];
            const int qa = (pa + W.nc[0]) % W.nc[0], qb = (pb + W.nc[1]) % W.nc[1];
            const int qc = (pc + W.nc[2]) % W.nc[2];
            const int cid = W.cell_off + (qa * W.nc[1] + qb) * W.nc[2] + qc;
            st = b.cell_start[cid];
            v = b.cell_count[cid];
        }
        int sum = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(sum, o);
            if (t >= o) sum += u;
        }
        if (t < ncell) {
            c_start[t] = st;
            c_pref[t + 1] = sum;
        }
        if (t == 0) {
            c_pref[0] = 0;
            n_cells = ncell;
        }
    }
    // the home atoms' box: their minimum-image offsets from the item's first
    // atom, as a centre and half widths per axis (an arc of the periodic cell
    // along each axis that holds every home atom, whatever the cell size)
    const int slot = threadIdx.x / kNbLanes;
    const int i0 = b.cell_atoms[s0 + chunk * kNbSlots];
    if (have && threadIdx.x % kNbLanes == 0) {
        double dx = xi - b.x[i0], dy = yi - b.y[i0], dz = zi - b.z[i0];
        min_image(dx, dy, dz, W);
        box[6 + 3 * slot] = dx;
        box[6 + 3 * slot + 1] = dy;
        box[6 + 3 * slot + 2] = dz;
    }
    __syncthreads();
    const int ncand = c_pref[n_cells];
    if (threadIdx.x < 3) {
        const int a = threadIdx.x, ns = min(kNbSlots, n0 - chunk * kNbSlots);
        double lo = box[6 + a], hi = lo;
        for (int k = 1; k < ns; ++k) {
            lo = fmin(lo, box[6 + 3 * k + a]);
            hi = fmax(hi, box[6 + 3 * k + a]);
        }
        box[a] = (a == 0 ? b.x[i0] : a == 1 ? b.y[i0] : b.z[i0]) + 0.5 * (lo + hi);
        box[3 + a] = 0.5 * (hi - lo);
    }
    __syncthreads();
    const double bcx = box[0], bcy = box[1], bcz = box[2], bhx = box[3], bhy = box[4], bhz = box[5];
    const double rl2 = b.field->rlist2;
    const double rn2 = b.field->rnear2;
    const double rz2 = b.field->rzone2;
    const int lane = threadIdx.x % kNbLanes;
    const int wlane = threadIdx.x & 63;
    static_assert(kNbLanes == 16, "the hit ballot is split in 16-bit fields");
    // a free atom of a zoned walker puts itself and every atom of its list
    // within rc + skin/2 in the walker's zone
    const bool mark = have && b.zflag && (b.zflag[w] & AKMC_ZONE_ON) && b.zfrozen[i] == 0;
    if (mark && lane == 0) b.zmark[i] = 1;
    int cnt = 0, nnear = 0;
    int ms = 0;  // candidates staged in LDS
    const int wave = threadIdx.x / 64;
    // pass 0 lists the near shell (r^2 <= rnear2), pass 1 the rest
    for (int pass = (rn2 > 0.0 ? 0 : 1); pass < 2; ++pass) {
        for (int seg = 0; seg < ncand; seg += kNbCap) {
            const int m = min(kNbCap, ncand - seg);
            if (!(pass == 1 && m == ncand && rn2 > 0.0)) {  // a single segment stays staged
                // every thread's candidates: cell by binary search over the
                // scan, index, then coordinates, each level's loads in flight
                // together before the LDS stores
                int jv[kNbStage];
                double vx[kNbStage], vy[kNbStage], vz[kNbStage];
#pragma unroll
                for (int u = 0; u < kNbStage; ++u) {
                    const int t = threadIdx.x + u * kBlock;
                    jv[u] = -1;
                    if (t < m) {
                        const int qq = seg + t;
                        int c = 0, hi = n_cells - 1;  // the last cell c with c_pref[c] <= qq
                        while (c < hi) {
                            const int mid = (c + hi + 1) >> 1;
                            if (c_pref[mid] <= qq) c = mid;
                            else hi = mid - 1;
                        }
                        jv[u] = b.cell_atoms[c_start[c] + qq - c_pref[c]];
                    }
                }
#pragma unroll
                for (int u = 0; u < kNbStage; ++u) {
                    if (jv[u] >= 0) {
                        vx[u] = b.x[jv[u]];
                        vy[u] = b.y[jv[u]];
                        vz[u] = b.z[jv[u]];
                    }
                }
                // only candidates within the list radius of the home atoms'
                // box are staged, compacted in candidate order; the distance
                // to the box per axis is the candidate's minimum-image offset
                // from the centre less the half width, never more than its
                // offset from any home atom, so every neighbour stays and
                // the lists are unchanged
                bool keep[kNbStage];
                int rank[kNbStage];
#pragma unroll
                for (int u = 0; u < kNbStage; ++u) {
                    keep[u] = false;
                    if (jv[u] >= 0) {
                        double dx = vx[u] - bcx, dy = vy[u] - bcy, dz = vz[u] - bcz;
                        min_image(dx, dy, dz, W);
                        const double ex = fmax(fabs(dx) - bhx, 0.0), ey = fmax(fabs(dy) - bhy, 0.0),
                                     ez = fmax(fabs(dz) - bhz, 0.0);
                        keep[u] = ex * ex + ey * ey + ez * ez <= rl2 * (1.0 + 1.0e-9);
                    }
                    const unsigned long long bal = __ballot(keep[u]);
                    rank[u] = __popcll(bal & ((1ull << wlane) - 1ull));
                    if (wlane == 0) scnt[u * kNbWaves + wave] = __popcll(bal);
                }
                __syncthreads();
                // a kept candidate's slot: the kept ones of earlier rounds,
                // of earlier wavefronts in its round, and of lower lanes
                int total = 0;
#pragma unroll
                for (int u = 0; u < kNbStage; ++u) {
                    for (int v = 0; v < kNbWaves; ++v) {
                        const int c = scnt[u * kNbWaves + v];
                        if (v < wave) rank[u] += c;
                        total += c;
                    }
                    if (u + 1 < kNbStage) rank[u + 1] += total;
                }
                ms = total;
#pragma unroll
                for (int u = 0; u < kNbStage; ++u) {
                    if (keep[u]) {
                        const int pos = rank[u];
                        sj[pos] = jv[u];
                        sx[pos] = vx[u];
                        sy[pos] = vy[u];
                        sz[pos] = vz[u];
                    }
                }
                __syncthreads();
            }
            for (int g = 0; g < ms; g += kNbLanes * kNbIlp) {
                int hit[kNbIlp], jj[kNbIlp];
                bool zin[kNbIlp];
#pragma unroll
                for (int u = 0; u < kNbIlp; ++u) {
                    const int t = g + kNbLanes * u + lane;
                    hit[u] = 0;
                    jj[u] = 0;
                    zin[u] = false;
                    if (have && t < ms) {
                        jj[u] = sj[t];
                        double dx = xi - sx[t], dy = yi - sy[t], dz = zi - sz[t];
                        min_image(dx, dy, dz, W);
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        const bool near = r2 <= rn2;
                        hit[u] = (jj[u] != i && r2 <= rl2 && (pass == 0 ? near : !near)) ? 1 : 0;
                        zin[u] = r2 <= rz2;
                    }
                }
#pragma unroll
                for (int u = 0; u < kNbIlp; ++u) {
                    // the atom's lanes are a 16-bit field of the wavefront's hit
                    // ballot: a lane's slot is the count of hits below it
                    const unsigned long long bal = __ballot(hit[u]);
                    const unsigned grp = unsigned(bal >> (wlane & ~(kNbLanes - 1))) & 0xFFFFu;
                    if (hit[u]) {
                        put_nbr(b, i, cnt + __popc(grp & ((1u << lane) - 1u)), uint32_t(jj[u] - W.off) | AKMC_NB_ZERO_SHIFT);
                        if (mark && zin[u]) b.zmark[jj[u]] = 1;
                    }
                    cnt += __popc(grp);
                }
            }
        }
        if (pass == 0) nnear = cnt;
    }
    if (have && lane == 0) {
        const int n = cnt < b.maxnb ? cnt : b.maxnb;
        b.nnb[i] = n;
        b.nnear[i] = nnear < b.maxnb ? nnear : b.maxnb;
        put_sentinel(b, i, n, W.off);
    }
}

// A fixed grid works through the items that k_cell_bin queued this step, so
// steps without rebuilds cost one short launch and the LDS of the staging
// buffers is claimed only by workgroups with work.  An item is encoded as
// (cell << 8) | chunk; chunks beyond 255 of very full cells are taken by the
// same item in strides of 256.
constexpr int kNbGrid = 1024;

__global__ void __launch_bounds__(kBlock) k_nbr_build_cells(BatchDev b)
{
    __shared__ double sx[kNbCap], sy[kNbCap], sz[kNbCap];
    __shared__ int sj[kNbCap];
    __shared__ int c_start[27], c_pref[28];
    __shared__ int n_cells;
    __shared__ double box[6 + 3 * kNbSlots];
    __shared__ int scnt[kNbStage * kNbWaves];
    const int items = *b.cell_work_n;
    for (int it = blockIdx.x; it < items; it += gridDim.x) {
        const int code = b.cell_work[it];
        const int cg = code >> 8, chunk0 = code & 255;
        const int n0 = b.cell_count[cg];
        for (int chunk = chunk0; chunk * kNbSlots < n0; chunk += 256) {
            nbr_build_item(b, cg, chunk, sx, sy, sz, sj, c_start, c_pref, n_cells, box, scnt);
            __syncthreads();
        }
    }
}

// Explicit-image lists (fields without the image convention) hold every
// (j, S) with |r_i - r_j - S L| <= rc + skin except j == i, S == 0; the shell
// loop is centred on the nearest image so unwrapped positions are handled.
// image_hits: one candidate j of atom i over those shells; calls put(entry)
// for each image inside the list radius and on the pass's side of the near
// shell, in (sa, sb, sc) ascending order
template <typename Put>
__device__ inline void image_hits(const BatchDev &b, const WalkerDev &W, int i, int jl, double xi, double yi,
                                  double zi, double rl2, double rn2, int pass, Put put)
{
    const int j = W.off + jl;
    const double bx = xi - b.x[j], by = yi - b.y[j], bz = zi - b.z[j];
    const int ma = (int)rint(bx * W.invL[0]), mb = (int)rint(by * W.invL[1]), mc = (int)rint(bz * W.invL[2]);
    for (int sa = ma - W.shells[0]; sa <= ma + W.shells[0]; ++sa) {
        const double dx = bx - sa * W.L[0];
        for (int sb = mb - W.shells[1]; sb <= mb + W.shells[1]; ++sb) {
            const double dy = by - sb * W.L[1];
            for (int sc = mc - W.shells[2]; sc <= mc + W.shells[2]; ++sc) {
                if (j == i && sa == 0 && sb == 0 && sc == 0) continue;
                const double dz = bz - sc * W.L[2];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const bool near = r2 <= rn2;
                if (r2 <= rl2 && (pass == 0 ? near : !near)) {
                    if (sa < -7 || sa > 7 || sb < -7 || sb > 7 || sc < -7 || sc > 7) {
                        *b.overflow = 1;
                        continue;
                    }
                    put(uint32_t(jl) | (uint32_t(sa + 7) << 20) | (uint32_t(sb + 7) << 24) | (uint32_t(sc + 7) << 28));
                }
            }
        }
    }
}

// Explicit-image list of a rebuilding walker (no-image fields, rc > L/2):
// one wavefront per atom.  Candidates are taken 64 at a time, one per
// lane; each lane counts its hits, a wavefront prefix sum gives every lane
// its first slot, and the lanes write their entries there.  The list holds
// the near-shell entries first, then the rest, each in (j, image)
// ascending order -- the order of a serial scan, so the list does not
// depend on scheduling.
__global__ void __launch_bounds__(kBlock) k_nbr_build_images(BatchDev b)
{
    constexpr int kAtomsPerBlock = kBlock / 64;
    const int nblk = (b.A + kAtomsPerBlock - 1) / kAtomsPerBlock;
    const int i = blockIdx.x * kAtomsPerBlock + int(threadIdx.x) / 64, lane = int(threadIdx.x) % 64;
    if (blockIdx.x >= nblk || i >= b.A) return;  // whole wavefronts leave together
    const int w = walker_of(b, i);
    if (!b.rebuild[w]) return;
    const WalkerDev W = b.walkers[w];
    const double rl2 = b.field->rlist2, rn2 = b.field->rnear2;
    const double xi = b.x[i], yi = b.y[i], zi = b.z[i];
    int cnt = 0, nnear = 0;
    for (int pass = (rn2 > 0.0 ? 0 : 1); pass < 2; ++pass) {
        for (int jb = 0; jb < W.n; jb += 64) {
            const int jl = jb + lane;
            int c = 0;
            if (jl < W.n) image_hits(b, W, i, jl, xi, yi, zi, rl2, rn2, pass, [&](uint32_t) { ++c; });
            int incl = c;
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(incl, o);
                if (lane >= o) incl += t;
            }
            int slot = cnt + incl - c;
            if (c > 0) image_hits(b, W, i, jl, xi, yi, zi, rl2, rn2, pass, [&](uint32_t e) { put_nbr(b, i, slot++, e); });
            cnt += __shfl(incl, 63);
        }
        if (pass == 0) nnear = cnt;
    }
    if (lane != 0) return;
    const int n = cnt < b.maxnb ? cnt : b.maxnb;
    b.nnb[i] = n;
    b.nnear[i] = nnear < b.maxnb ? nnear : b.maxnb;
    put_sentinel(b, i, n, W.off);
}

// The active zone of a search walker whose lists were rebuilt (one
// workgroup per walker): the list build has marked its free atoms and every
// atom on a free atom's list (on top of the marks since the walker's start,
// cleared by k_cell_bin at a restart); the members are listed in ascending
// index order, each thread a contiguous run of the walker's atoms.
__global__ void __launch_bounds__(kBlock) k_zone_build(BatchDev b, unsigned long long *count)
{
    __shared__ int part[kBlock / 64];
    const int w = blockIdx.x;
    if (!b.rebuild[w]) return;
    const int f = b.zflag[w];
    // the walker's rebuild is done (k_rebuild_done's bookkeeping)
    if (!(f & AKMC_ZONE_ON)) {
        if (threadIdx.x == 0) {
            b.rebuild[w] = 0;
            atomicAdd(count, 1ull);
        }
        return;
    }
    const WalkerDev W = b.walkers[w];
    const int per = (W.n + kBlock - 1) / kBlock;
    const int l0 = (int)threadIdx.x * per, l1 = min(W.n, l0 + per);
    // the run's marks, kept as bits (the first 32) so the listing below
    // stores without reloading them
    int c = 0;
    unsigned bits = 0;
    for (int l = l0; l < l1; ++l) {
        const bool m = b.zmark[W.off + l] != 0;
        c += m;
        if (m && l - l0 < 32) bits |= 1u << (l - l0);
    }
    // inclusive scan of the runs' counts: wavefront shuffles, then the
    // wavefront totals in order
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = c;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) part[wave] = incl;
    __syncthreads();
    for (int k = 0; k < wave; ++k) incl += part[k];
    int at = incl - c;
    for (int l = l0; l < l1; ++l)
        if (l - l0 < 32 ? ((bits >> (l - l0)) & 1u) != 0 : b.zmark[W.off + l] != 0) b.zlist[W.off + at++] = l;
    if (threadIdx.x == kBlock - 1) {
        b.nz[w] = incl;
        b.zflag[w] = f & ~AKMC_ZONE_RESET;
        b.rebuild[w] = 0;
        atomicAdd(count, 1ull);
    }
}

// The work lists of a search micro-step (one 1024-thread workgroup, after
// the zones): for each list li (FP32 pass 1, FP32 pass 2, FP64 pass 1,
// FP64 pass 2) with chunk size apc[li] > 0, the chunks of every walker in
// that precision, walker after walker, each chunk apc[li] consecutive slots
// of the walker's pass list.
__global__ void __launch_bounds__(1024) k_work_list(BatchDev b, int4 apc)
{
    __shared__ int wsum[16];
    __shared__ int carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        const int li = blockIdx.x;  // one workgroup per list
        const int a = li == 0 ? apc.x : li == 1 ? apc.y : li == 2 ? apc.z : apc.w;
        if (tid == 0) carry = 0;
        __syncthreads();
        if (a > 0) {
            const int mode = li < 2 ? 1 : 2, pass = (li & 1) + 1;
            const int full = mode == 2 ? AKMC_ZONE_FULL64 : AKMC_ZONE_FULL32;
            for (int w0 = 0; w0 < b.W; w0 += 1024) {
                const int w = w0 + tid;
                int k = 0;
                if (w < b.W && b.active[w] == mode) {
                    int len = b.walkers[w].n;
                    if (b.zflag) {
                        const int f = b.zflag[w];
                        if ((f & AKMC_ZONE_ON) && !(f & full)) len = pass == 1 ? b.nz[w] : b.znact[w];
                    }
                    k = (len + a - 1) / a;
                }
                int incl = k;
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(incl, o);
                    if (lane >= o) incl += t;
                }
                if (lane == 63) wsum[wave] = incl;
                __syncthreads();
                int at = carry + incl - k;
                for (int v = 0; v < wave; ++v) at += wsum[v];
                for (int j = 0; j < k; ++j) b.wl_tab[li][at + j] = make_int2(w, j * a);
                __syncthreads();
                if (tid == 0)
                    for (int v = 0; v < 16; ++v) carry += wsum[v];
                __syncthreads();
            }
        }
        if (tid == 0) b.wl_count[li] = carry;
        __syncthreads();
    }
}

__global__ void k_rebuild_done(BatchDev b, unsigned long long *count)
{
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= b.W) return;
    if (b.rebuild[w]) {
        b.rebuild[w] = 0;
        atomicAdd(count, 1ull);
    }
}

// ---------------------------------------------------------- Sutton-Chen

// the global thread index in XCD order (see xcd_atom), -1 past the end
__device__ inline int xcd_thread(int nblocks)
{
    const int per = (nblocks + kXCD - 1) / kXCD;
    const int logical = (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
    return logical < nblocks ? logical * blockDim.x + threadIdx.x : -1;
}

// sum over the LPA consecutive lanes that share one atom
template <int LPA>
__device__ inline double lane_group_sum(double v)
{
    for (int o = LPA / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// The Sutton-Chen pair table (nspec^2 entries) staged in LDS by every
// workgroup before any thread leaves, so the coefficient lookup of a pair is
// an LDS read instead of a third dependent global load.
__device__ inline void stage_sc_table(const FieldDev &F, SCPairDev *tab)
{
    static_assert(sizeof(SCPairDev) % 4 == 0, "table staged in 4-byte words");
    const int nw = F.nspec * F.nspec * int(sizeof(SCPairDev) / 4);
    const uint32_t *src = reinterpret_cast<const uint32_t *>(F.sc);
    uint32_t *dst = reinterpret_cast<uint32_t *>(tab);
    for (int w = threadIdx.x; w < nw; w += blockDim.x) dst[w] = src[w];
    __syncthreads();
}


// The neighbour loops below hold U list entries, fetch the next U entries
// while gathering the current U neighbours' positions (and densities), then
// evaluate the U pairs: U independent gathers per lane in flight together
// with the next batch of the (streamed) list, instead of one dependent
// chain per pair.  The
// per-lane summation order is k ascending for every U (the FP32 variants
// agree bitwise; FP64 ones to the last bit of FMA contraction); LPA variants
// differ by summation order.

// list entries k, k + LPA, ..., k + (U-1) LPA of atom i.  A slot past the
// end reads row cnt, the list's sentinel (i's own entry, rejected by the
// r^2 floor), so the loads are unconditional: no branch separates them from
// the batch they prefetch for, and the wait before a batch's first use
// counts only the loads issued before it.  NB16: the batch is known to store
// 16-bit minimum-image entries, read without the run-time width test (their
// image bits are never used).
template <int LPA, int U, bool NB16>
__device__ inline void fetch_entries(const BatchDev &b, int i, int k, int cnt, uint32_t (&e)[U])
{
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int kk = k + u * LPA;
        const int row = kk < cnt ? kk : cnt;
        // element offset row A + i in 32 bits, one 24-bit multiply-add (the
        // batch keeps A below 2^24 and the list's (maxnb + pad) A below 2^32)
        const uint32_t o = __umul24(uint32_t(row), uint32_t(b.A)) + uint32_t(i);
        if constexpr (NB16) e[u] = reinterpret_cast<const uint16_t *>(b.nbr)[o];
        else e[u] = nbr_entry(b, row, i);
    }
}

// A neighbour's packed gather record, whole (W4) or its position only
// (24 bytes: one 16-byte and one 8-byte read, w = 0)
template <bool W4>
__device__ inline double4 gather_rec(const double4 *base, int j)
{
    if constexpr (W4) {
        return base[j];
    } else {
        const double2 xy = reinterpret_cast<const double2 *>(base)[2 * j];
        const double z = reinterpret_cast<const double *>(base)[4 * j + 2];
        return make_double4(xy.x, xy.y, z, 0.0);
    }
}

// 1/sqrt(rho) of an FP32 walker in the fourth word of its packed gather
// record (single pair type): the float's bits in the low half, so the pass-2
// gather reads it without a conversion
__device__ inline double isr32_word(float v) { return __longlong_as_double((long long)__float_as_uint(v)); }
__device__ inline float isr32_of(double w) { return __uint_as_float(uint32_t(__double_as_longlong(w))); }

// The atom a thread serves.  Unstaged launches: LPA consecutive threads per
// atom over the whole batch (XCD-ordered).  Staged launches: one workgroup of
// AKMC_STAGE_BLOCK threads per chunk of AKMC_STAGE_BLOCK / LPA atoms of one
// walker; the workgroup first copies the walker's packed gather records
// (p4) into LDS, so every neighbour gather of the pair loop is an LDS read
// instead of a texture-unit gather from L2.  A workgroup whose walker runs
// in the other precision, or is inactive, leaves before staging.
// Work lists of the search micro-steps (k_work_list): item c of list li is
// (walker, first slot) of a chunk of the walker's pass list -- its zone
// (pass 1) or free atoms (pass 2) when zoned, else all its atoms.  The pair
// kernels of a search micro-step run a persistent grid over the list, each
// workgroup a contiguous range of it.  The grid is a multiple of kXCD.
__device__ inline int wl_first()
{
    const int per = gridDim.x / kXCD;
    return (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
}
// the workgroup's contiguous range [x, y) of a list of cnt items: equal
// shares in XCD order, so one walker's consecutive chunks mostly fall to one
// workgroup (one staging) and neighbouring ranges to one XCD
__device__ inline int2 wl_range(int cnt)
{
    const int L = wl_first(), G = gridDim.x;
    const int per = cnt / G, extra = cnt % G;
    const int x = L * per + min(L, extra);
    return make_int2(x, x + per + (L < extra ? 1 : 0));
}
template <typename T>
__device__ inline int wl_len(const BatchDev &b, const WalkerDev &W, int w, int pass)
{
    return zoned<T>(b, w) ? (pass == 1 ? b.nz[w] : b.znact[w]) : W.n;
}
template <typename T>
__device__ inline int wl_atom(const BatchDev &b, const WalkerDev &W, int w, int k, int pass)
{
    return zoned<T>(b, w) ? (pass == 1 ? W.off + b.zlist[W.off + k] : b.zact[W.off + k]) : W.off + k;
}

// With zp (1: pass 1, 2: pass 2) a zoned walker's slots run over its zone
// (pass 1) or its free atoms (pass 2) instead of all its atoms: slot k of
// the walker serves the k-th listed atom, and workgroups whose slots all
// lie past the list leave before staging.
struct AtomSlot {
    int i, sub;
    bool ok;
};

__device__ inline int zone_count(const BatchDev &b, int w, int zp) { return zp == 1 ? b.nz[w] : b.znact[w]; }
__device__ inline int zone_atom(const BatchDev &b, const WalkerDev &W, int k, int zp)
{
    return zp == 1 ? W.off + b.zlist[W.off + k] : b.zact[W.off + k];
}

template <typename T, int LPA, bool STAGED>
__device__ inline AtomSlot sc_slot(const BatchDev &b, WalkerDev &W, double4 *stage, int zp = 0)
{
    if constexpr (STAGED) {
        constexpr int li = LPA == 1 ? 0 : LPA == 2 ? 1 : LPA == 4 ? 2 : 3;
        const int nch = b.nchunk_tab[li];
        const int per = (nch + kXCD - 1) / kXCD;
        const int c = (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
        if (c >= nch) return {0, 0, false};
        const int2 ch = b.chunk_tab[li][c];
        if (other_precision<T>(b, ch.x)) return {0, 0, false};
        const bool z = zp && zoned<T>(b, ch.x);
        const int cnt = z ? zone_count(b, ch.x, zp) : 0;
        if (z && ch.y >= cnt) return {0, 0, false};
        W = b.walkers[ch.x];
        for (int k = threadIdx.x; k < W.n; k += blockDim.x) stage[k] = b.p4[W.off + k];
        __syncthreads();
        const int la = ch.y + int(threadIdx.x) / LPA;
        if (z) return {la < cnt ? zone_atom(b, W, la, zp) : 0, int(threadIdx.x) % LPA, la < cnt};
        return {W.off + la, int(threadIdx.x) % LPA, la < W.n};
    } else {
        const int t = xcd_thread((b.A * LPA + kBlock - 1) / kBlock);
        if (t < 0) return {0, 0, false};
        const int p = t / LPA;
        if (p >= b.A) return {0, 0, false};
        const int w = walker_of(b, p);
        if (other_precision<T>(b, w)) return {0, 0, false};
        W = b.walkers[w];
        if (zp && zoned<T>(b, w)) {
            const int k = p - W.off;
            if (k >= zone_count(b, w, zp)) return {0, 0, false};
            return {zone_atom(b, W, k, zp), t % LPA, true};
        }
        return {p, t % LPA, true};
    }
}

// pass 1: rho_i = sum_j (eps c)^2 (a/r)^m (potentials/metal/Field.cpp:1782-1853),
// then F(rho) = -sqrt(rho) (ManyBody.cpp:554-575) and 1/sqrt(rho) for pass 2.
// LPA lanes share one atom's neighbour list (strided) and combine by shuffles.
// ONE: the field has a single species pair with n = 9, m = 6, whose
// coefficients stay in registers and whose powers are straight products.
// ZONE (search micro-steps): zoned walkers visit their zone, and every
// visited atom also gets its half of its pairs' V (pass 2 then writes no
// energy, so the energies do not depend on pass 2's variant, and the
// frozen zone atoms' pair energies stay current while pass 2 visits only
// the free atoms).
template <typename T, bool MINIMG, int LPA, int U, bool ONE, bool STAGED, bool ZONE>
__device__ __forceinline__ void sc_density_atom(const BatchDev &b, const FieldDev &F, const SCPairDev *tab,
                                                const SCCoef<T> &c1, const WalkerDev &W, const double4 *stage,
                                                int i, int sub)
{
    constexpr bool NB16 = STAGED && MINIMG;  // staged batches store 16-bit minimum-image entries
    const double xi = b.x[i], yi = b.y[i], zi = b.z[i];
    const int row = b.type[i] * F.nspec;
    const double rc2 = F.rc2;
    const int cnt = b.nnb[i];
    double rho = 0.0, ep = 0.0;
    // entries two batches ahead: the next batch and the one after it are in
    // flight while a batch's gathers and pair terms run (more bytes of the
    // streamed list outstanding per wavefront)
    uint32_t en[U], en2[U];
    fetch_entries<LPA, U, NB16>(b, i, sub, cnt, en);
    fetch_entries<LPA, U, NB16>(b, i, sub + LPA * U, cnt, en2);
    for (int k0 = sub; k0 < cnt; k0 += LPA * U) {
        uint32_t e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            e[u] = en[u];
            en[u] = en2[u];
        }
        // the batch after next in flight; in the staged kernels (gathers from
        // LDS) a scheduling barrier keeps these loads here, ahead of this
        // batch's work, instead of next to their use two batches on
        fetch_entries<LPA, U, NB16>(b, i, k0 + 2 * LPA * U, cnt, en2);
        if constexpr (STAGED) __builtin_amdgcn_sched_barrier(0);
        double4 pj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int jl = NB16 ? int(e[u]) : int(e[u] & AKMC_NB_JMASK);
            pj[u] = STAGED ? gather_rec<!ONE>(stage, jl) : gather_rec<!ONE>(b.p4, W.off + jl);
        }
        double r2[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double dx = xi - pj[u].x, dy = yi - pj[u].y, dz = zi - pj[u].z;
            entry_disp(e[u], dx, dy, dz, W, MINIMG);
            r2[u] = dx * dx + dy * dy + dz * dz;
            in[u] = r2[u] <= rc2 && r2[u] > (MINIMG ? 1.0e-16 : 1.0e-3);
        }
        if constexpr (ONE) {
            // every slot is evaluated (padding gives r2 = 0 and infinities),
            // only the pairs inside the cutoff are summed
            T ir2[U], u2[U], rt[U], um[U];
            sc_powers_in<T, U>(r2, c1.a, ir2, u2, rt);
            if constexpr (ZONE) {
                T un[U];
                pow_9_6<T, U>(u2, rt, un, um);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    T eh = in[u] ? T(0.5) * c1.eps * un[u] : T(0);
                    if constexpr (sizeof(T) == 4) asm("" : "+v"(eh));
                    ep += double(eh);
                }
            } else {
                pow_6<T, U>(u2, um);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) rho += double(in[u] ? c1.rho_coef * um[u] : T(0));  // + exact zeros
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!in[u]) continue;
                const SCPairDev &p = tab[row + (int)pj[u].w];
                if (p.on) {
                    const SCCoef<T> &c = sc_coef<T>(p);
                    T u2 = c.a * c.a * recip<T>(T(r2[u]));
                    rho += double(c.rho_coef * pow_half<T>(u2, p.m_int, c.m));
                    if constexpr (ZONE) ep += double(T(0.5) * c.eps * pow_half<T>(u2, p.n_int, c.n));
                }
            }
        }
    }
    rho = lane_group_sum<LPA>(rho);
    if constexpr (ZONE) ep = lane_group_sum<LPA>(ep);
    if (sub != 0) return;
    b.rho[i] = rho;
    energy_a<T>(b)[i] = -sqrt(rho);
    if constexpr (ZONE) energy_b<T>(b)[i] = ep;
    const float isr32 = rho < 1.0e-8 ? -1.0f : float(1.0 / sqrt(rho));
    const double isr64 = rho < 1.0e-8 ? -1.0 : 1.0 / sqrt(rho);
    b.inv_sqrt_rho[i] = isr32;
    if (sizeof(T) == 8) b.isr64[i] = isr64;
    // single pair type: the packed gather record carries 1/sqrt(rho) in place
    // of the species for pass 2 (k_pack_check restores the species every step)
    if (ONE) reinterpret_cast<double *>(&b.p4[i])[3] = sizeof(T) == 8 ? isr64 : isr32_word(isr32);
}

template <typename T, bool MINIMG, int LPA, int U, bool ONE, bool STAGED, bool ZONE>
__global__ void __launch_bounds__(STAGED ? AKMC_STAGE_BLOCK : kBlock) k_sc_density(BatchDev b)
{
    __shared__ SCPairDev tab[ONE ? 1 : AKMC_MAXPAIR];
    extern __shared__ double4 stage[];
    const FieldDev &F = *b.field;
    if (!ONE) stage_sc_table(F, tab);
    const SCCoef<T> c1 = sc_coef<T>(F.sc[0]);
    WalkerDev W;
    if constexpr (ZONE) {
        // search micro-steps: a persistent grid over the step's work list
        const int li = sizeof(T) == 8 ? 2 : 0;
        const int2 rg = wl_range(b.wl_count[li]);
        int staged_w = -1;
        for (int c = rg.x; c < rg.y; ++c) {
            const int2 it = b.wl_tab[li][c];
            W = b.walkers[it.x];
            const int n = wl_len<T>(b, W, it.x, 1);
            if constexpr (STAGED) {
                if (it.x != staged_w) {  // a walker's chunks are consecutive: stage it once
                    __syncthreads();     // the previous items are done with the stage
                    for (int k = threadIdx.x; k < W.n; k += blockDim.x) stage[k] = b.p4[W.off + k];
                    __syncthreads();
                    staged_w = it.x;
                }
            }
            const int la = it.y + (int)threadIdx.x / LPA;
            if (la < n)
                sc_density_atom<T, MINIMG, LPA, U, ONE, STAGED, ZONE>(b, F, tab, c1, W, stage,
                                                                       wl_atom<T>(b, W, it.x, la, 1),
                                                                       (int)threadIdx.x % LPA);
        }
        return;
    }
    const AtomSlot sl = sc_slot<T, LPA, STAGED>(b, W, stage, 0);
    if (!sl.ok) return;
    sc_density_atom<T, MINIMG, LPA, U, ONE, STAGED, ZONE>(b, F, tab, c1, W, stage, sl.i, sl.sub);
}

// pass 2: F_i = sum_j r_ij (fs - fm)/r^2, fs = n eps (a/r)^n,
// fm = 1/2 m (eps c)^2 (1/sqrt(rho_i) + 1/sqrt(rho_j)) (a/r)^m
// (potentials/metal/Field.cpp:587-640, ManyBody.cpp:621-757); each atom
// also carries half of each pair's V = eps (a/r)^n
template <typename T, bool MINIMG, bool STRESS, int LPA, int U, bool ONE, bool STAGED>
__device__ __forceinline__ void sc_force_atom(const BatchDev &b, const FieldDev &F, const SCPairDev *tab,
                                              const SCCoef<T> &c1, const WalkerDev &W, const double4 *stage,
                                              int i, int sub)
{
    constexpr bool NB16 = STAGED && MINIMG;
    const double xi = b.x[i], yi = b.y[i], zi = b.z[i];
    const int row = b.type[i] * F.nspec;
    const double rc2 = F.rc2;
    const int cnt = b.nnb[i];
    const float isi = sizeof(T) == 8 ? 0.0f : b.inv_sqrt_rho[i];
    const double isi64 = sizeof(T) == 8 ? b.isr64[i] : 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    // entries two batches ahead: the next batch and the one after it are in
    // flight while a batch's gathers and pair terms run (more bytes of the
    // streamed list outstanding per wavefront)
    uint32_t en[U], en2[U];
    fetch_entries<LPA, U, NB16>(b, i, sub, cnt, en);
    fetch_entries<LPA, U, NB16>(b, i, sub + LPA * U, cnt, en2);
    for (int k0 = sub; k0 < cnt; k0 += LPA * U) {
        uint32_t e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            e[u] = en[u];
            en[u] = en2[u];
        }
        // the batch after next in flight; in the staged kernels (gathers from
        // LDS) a scheduling barrier keeps these loads here, ahead of this
        // batch's work, instead of next to their use two batches on
        fetch_entries<LPA, U, NB16>(b, i, k0 + 2 * LPA * U, cnt, en2);
        if constexpr (STAGED) __builtin_amdgcn_sched_barrier(0);
        double4 pj[U];
        float isj[U];
        double isj64[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int jl = NB16 ? int(e[u]) : int(e[u] & AKMC_NB_JMASK);
            const int j = W.off + jl;
            pj[u] = STAGED ? stage[jl] : b.p4[j];
            if (ONE) {  // 1/sqrt(rho_j) rides in the gather record (pass 1)
                if (sizeof(T) == 8) isj64[u] = pj[u].w;
                else isj[u] = isr32_of(pj[u].w);
            } else {
                if (sizeof(T) == 8) isj64[u] = b.isr64[j];
                else isj[u] = b.inv_sqrt_rho[j];
            }
        }
        double dx[U], dy[U], dz[U], r2[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            dx[u] = xi - pj[u].x;
            dy[u] = yi - pj[u].y;
            dz[u] = zi - pj[u].z;
            entry_disp(e[u], dx[u], dy[u], dz[u], W, MINIMG);
            r2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
            in[u] = r2[u] <= rc2 && r2[u] > (MINIMG ? 1.0e-16 : 1.0e-3);
        }
        T ir2[U], un[U], um[U];
        if constexpr (ONE) {
            // every slot is evaluated (padding gives r2 = 0 and infinities),
            // only the pairs inside the cutoff are summed
            T u2[U], rt[U];
            sc_powers_in<T, U>(r2, c1.a, ir2, u2, rt);
            pow_9_6<T, U>(u2, rt, un, um);
        }
        if constexpr (ONE && !STRESS) {
            // branch-free: a pair outside the cutoff (or a padding slot)
            // adds exact zeros, which leaves the FP64 sums bitwise as if
            // it were skipped
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const T fs = c1.fs_coef * un[u];
                T fm;
                if constexpr (sizeof(T) == 8)
                    fm = (isi64 >= 0.0 && isj64[u] >= 0.0) ? c1.fm_coef * T(isi64 + isj64[u]) * um[u] : T(0);
                else
                    fm = (isi >= 0.0f && isj[u] >= 0.0f) ? c1.fm_coef * (T(isi) + T(isj[u])) * um[u] : T(0);
                T ft = in[u] ? (fs - fm) * ir2[u] : T(0);
                T eh = in[u] ? T(0.5) * c1.eps * un[u] : T(0);
                if constexpr (sizeof(T) == 4) {
                    // keep the selects on the FP32 values (one instruction
                    // each) instead of letting them sink past the widening
                    asm("" : "+v"(ft));
                    asm("" : "+v"(eh));
                }
                const double dft = double(ft);
                fx += dft * dx[u];
                fy += dft * dy[u];
                fz += dft * dz[u];
                ep += double(eh);
            }
            continue;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!in[u]) continue;
            const SCPairDev &p = ONE ? F.sc[0] : tab[row + (int)pj[u].w];
            if (!(ONE || p.on)) continue;
            const SCCoef<T> &c = ONE ? c1 : sc_coef<T>(p);
            if constexpr (!ONE) {
                ir2[u] = recip<T>(T(r2[u]));
                const T u2 = c.a * c.a * ir2[u];
                un[u] = pow_half<T>(u2, p.n_int, c.n);
                um[u] = pow_half<T>(u2, p.m_int, c.m);
            }
            T fs = c.fs_coef * un[u];
            T fm = T(0);
            if (sizeof(T) == 8) {
                if (isi64 >= 0.0 && isj64[u] >= 0.0) fm = c.fm_coef * T(isi64 + isj64[u]) * um[u];
            } else if (isi >= 0.0f && isj[u] >= 0.0f) {
                fm = c.fm_coef * (T(isi) + T(isj[u])) * um[u];
            }
            T ft = (fs - fm) * ir2[u];  // both orderings of the pair
            double dft = double(ft);
            fx += dft * dx[u];
            fy += dft * dy[u];
            fz += dft * dz[u];
            ep += double(T(0.5) * c.eps * un[u]);
            if (STRESS) {
                double h = -0.5 * dft;  // one ordered pair: -(fs-fm)/(2 r^2) r r
                s0 += h * dx[u] * dx[u];
                s1 += h * dy[u] * dy[u];
                s2 += h * dz[u] * dz[u];
                s3 += h * dy[u] * dz[u];
                s4 += h * dx[u] * dz[u];
                s5 += h * dx[u] * dy[u];
            }
        }
    }
    fx = lane_group_sum<LPA>(fx);
    fy = lane_group_sum<LPA>(fy);
    fz = lane_group_sum<LPA>(fz);
    ep = lane_group_sum<LPA>(ep);
    if (STRESS) {
        s0 = lane_group_sum<LPA>(s0);
        s1 = lane_group_sum<LPA>(s1);
        s2 = lane_group_sum<LPA>(s2);
        s3 = lane_group_sum<LPA>(s3);
        s4 = lane_group_sum<LPA>(s4);
        s5 = lane_group_sum<LPA>(s5);
    }
    if (sub != 0) return;
    b.fx[i] = fx;
    b.fy[i] = fy;
    b.fz[i] = fz;
    if (!b.pe1) energy_b<T>(b)[i] = ep;  // search micro-steps: pass 1 wrote it
    if (STRESS) {
        b.st_atom[0 * (size_t)b.A + i] = s0;
        b.st_atom[1 * (size_t)b.A + i] = s1;
        b.st_atom[2 * (size_t)b.A + i] = s2;
        b.st_atom[3 * (size_t)b.A + i] = s3;
        b.st_atom[4 * (size_t)b.A + i] = s4;
        b.st_atom[5 * (size_t)b.A + i] = s5;
    }
}

// WL: search micro-steps, a persistent grid over the step's work list
template <typename T, bool MINIMG, bool STRESS, int LPA, int U, bool ONE, bool STAGED, bool WL>
__global__ void __launch_bounds__(STAGED ? AKMC_STAGE_BLOCK : kBlock) k_sc_force(BatchDev b)
{
    __shared__ SCPairDev tab[ONE ? 1 : AKMC_MAXPAIR];
    extern __shared__ double4 stage[];
    const FieldDev &F = *b.field;
    if (!ONE) stage_sc_table(F, tab);
    const SCCoef<T> c1 = sc_coef<T>(F.sc[0]);
    WalkerDev W;
    if constexpr (WL) {
        const int li = sizeof(T) == 8 ? 3 : 1;
        const int2 rg = wl_range(b.wl_count[li]);
        int staged_w = -1;
        for (int c = rg.x; c < rg.y; ++c) {
            const int2 it = b.wl_tab[li][c];
            W = b.walkers[it.x];
            const int n = wl_len<T>(b, W, it.x, 2);
            if constexpr (STAGED) {
                if (it.x != staged_w) {  // a walker's chunks are consecutive: stage it once
                    __syncthreads();     // the previous items are done with the stage
                    for (int k = threadIdx.x; k < W.n; k += blockDim.x) stage[k] = b.p4[W.off + k];
                    __syncthreads();
                    staged_w = it.x;
                }
            }
            const int la = it.y + (int)threadIdx.x / LPA;
            if (la < n)
                sc_force_atom<T, MINIMG, STRESS, LPA, U, ONE, STAGED>(b, F, tab, c1, W, stage,
                                                                       wl_atom<T>(b, W, it.x, la, 2),
                                                                       (int)threadIdx.x % LPA);
        }
        return;
    }
    const AtomSlot sl = sc_slot<T, LPA, STAGED>(b, W, stage, 2);
    if (!sl.ok) return;
    sc_force_atom<T, MINIMG, STRESS, LPA, U, ONE, STAGED>(b, F, tab, c1, W, stage, sl.i, sl.sub);
}

// energy-only pass 2 of Sutton-Chen (Field::calculateEnergy,
// potentials/metal/Field.cpp:1088-1297): each atom's half of its pairs'
// V = eps (a/r)^n, no forces.  One lane per atom; species from the type
// array (the ONE kernels' gather records carry 1/sqrt(rho) in their fourth
// word after pass 1).
template <typename T, bool MINIMG>
__global__ void __launch_bounds__(kBlock) k_sc_pair_energy(BatchDev b)
{
    const int i = xcd_atom((b.A + kBlock - 1) / kBlock);
    if (i < 0 || i >= b.A) return;
    const int w = walker_of(b, i);
    if (other_precision<T>(b, w)) return;
    const FieldDev &F = *b.field;
    const WalkerDev W = b.walkers[w];
    const double xi = b.x[i], yi = b.y[i], zi = b.z[i];
    const int row = b.type[i] * F.nspec;
    const int cnt = b.nnb[i];
    double ep = 0.0;
    for (int k = 0; k < cnt; ++k) {
        const uint32_t e = nbr_entry(b, k, i);
        const int j = W.off + int(e & AKMC_NB_JMASK);
        const double4 pj = b.p4[j];
        double dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
        entry_disp(e, dx, dy, dz, W, MINIMG);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 <= F.rc2) || !(r2 > (MINIMG ? 1.0e-16 : 1.0e-3))) continue;
        const SCPairDev &p = F.sc[row + b.type[j]];
        if (!p.on) continue;
        const SCCoef<T> &c = sc_coef<T>(p);
        const T u2 = c.a * c.a * recip<T>(T(r2));
        ep += double(T(0.5) * c.eps * pow_half<T>(u2, p.n_int, c.n));
    }
    energy_b<T>(b)[i] = ep;
}

// ------------------------------------------------------- tabulated EAM

// eamMeshEnergy (potentials/metal/ManyBody.cpp:455-460) on a table stored as
// (t_i, t_{i+1} - t_i): t_i + ((t_{i+1} - t_i) (x - i h)) / h.  The index
// is clamped to the last interval; a density past its table is flagged by
// the caller (the reference exits there, and reads one entry past the table
// for the last index).
__device__ inline double eam_interp(const double2 *t, int stride, int n, double x, double h, int &idx)
{
    idx = int(x / h);
    const int k = idx < 0 ? 0 : (idx > n - 2 ? n - 2 : idx);
    const double2 v = t[(size_t)k * stride];
    return v.x + (v.y * (x - k * h)) / h;
}

constexpr double kEamTiny = 1.0e-8;  // ManyBody.cpp:728
constexpr double kEamNoDeriv = -1.0e300;  // F'(rho) of an atom whose rho < tiny

// pass 1: rho_i = sum_j rho_{t_j t_j}(r) over the pairs inside both the
// field cutoff and the table's (Field.cpp:1782-1853 with
// calculateManyBodyDensEnergy, ManyBody.cpp:462-545), then
// F_{t_i t_i}(rho_i) (embed, ManyBody.cpp:546-619) and F'(rho_i).  The
// interpolation runs in FP64 on both precision paths: a table lookup is a
// gather, not arithmetic, so the FP32 path gains nothing from FP32 here.
// ONE: one table for every pair; F'(rho_i) then rides in the packed gather
// record for pass 2 (k_pack_check restores the species every step).
template <typename T, bool MINIMG, int LPA, bool ONE, bool STAGED>
__global__ void __launch_bounds__(STAGED ? AKMC_STAGE_BLOCK : kBlock) k_eam_density(BatchDev b)
{
    extern __shared__ double4 stage[];
    const FieldDev &F = *b.field;
    WalkerDev W;
    const AtomSlot sl = sc_slot<T, LPA, STAGED>(b, W, stage);
    if (!sl.ok) return;
    const int i = sl.i, sub = sl.sub;
    const double xi = b.x[i], yi = b.y[i], zi = b.z[i];
    const double rc2 = F.rc2;
    const int cnt = b.nnb[i];
    const int ns = F.nspec;
    double rho = 0.0;
    for (int k = sub; k < cnt; k += LPA) {
        const uint32_t e = nbr_entry(b, k, i);
        const int jl = int(e & AKMC_NB_JMASK);
        const double4 pj = STAGED ? stage[jl] : b.p4[W.off + jl];
        double dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
        entry_disp(e, dx, dy, dz, W, MINIMG);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 <= rc2 && r2 > (MINIMG ? 1.0e-16 : 1.0e-3))) continue;
        const int tj = ONE ? 0 : (int)pj.w;
        const int tb = ONE ? F.eam_tab[0] : F.eam_tab[tj * ns + tj];  // density of species b
        if (tb < 0) continue;
        const EamTabDev &T_ = F.eamt[tb];
        const double r = sqrt(r2);
        if (r < kEamTiny || !(r < T_.cut)) continue;
        int idx;
        rho += eam_interp(T_.dens, 1, T_.nr, r, T_.dr, idx);
    }
    rho = lane_group_sum<LPA>(rho);
    if (sub != 0) return;
    b.rho[i] = rho;
    const int ti = ONE ? 0 : b.type[i];
    const int ta = ONE ? F.eam_tab[0] : F.eam_tab[ti * ns + ti];
    double Fi = 0.0, Fp = kEamNoDeriv;
    if (ta >= 0) {
        const EamTabDev &T_ = F.eamt[ta];
        int idx;
        Fi = eam_interp(T_.emb, 2, T_.nrho, rho, T_.drho, idx);
        if (idx < 0 || idx >= T_.nrho) atomicOr(b.overflow, AKMC_OV_EAM_RHO);
        if (rho >= kEamTiny) Fp = eam_interp(T_.emb + 1, 2, T_.nrho, rho, T_.drho, idx);
    }
    energy_a<T>(b)[i] = Fi;
    if (ONE) reinterpret_cast<double *>(&b.p4[i])[3] = Fp;
}

// pass 2: per ordered pair (Field.cpp:587-640 with
// calculateManyBodyPairForce and calculateManyBodyForce,
// ManyBody.cpp:621-835): fs = -r phi'(r), fm = (F'(rho_i) + F'(rho_j))
// rho'(r) r (zero when either density is below 1e-8), F_i += r_ij
// (fs - fm) / r^2 over both orderings of each pair, and half of phi(r) to
// the pair energy.  phi, phi' and rho' are those of the (t_i, t_j) table,
// F' too, evaluated at rho_i and rho_j, as the reference does.
template <typename T, bool MINIMG, bool STRESS, int LPA, bool ONE, bool STAGED>
__global__ void __launch_bounds__(STAGED ? AKMC_STAGE_BLOCK : kBlock) k_eam_force(BatchDev b)
{
    extern __shared__ double4 stage[];
    const FieldDev &F = *b.field;
    WalkerDev W;
    const AtomSlot sl = sc_slot<T, LPA, STAGED>(b, W, stage);
    if (!sl.ok) return;
    const int i = sl.i, sub = sl.sub;
    const double xi = b.x[i], yi = b.y[i], zi = b.z[i];
    const double rc2 = F.rc2;
    const int cnt = b.nnb[i];
    const int ns = F.nspec;
    const int ti = ONE ? 0 : b.type[i];
    const double rhoi = b.rho[i];
    const double fpi = ONE ? b.p4[i].w : 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    for (int k = sub; k < cnt; k += LPA) {
        const uint32_t e = nbr_entry(b, k, i);
        const int jl = int(e & AKMC_NB_JMASK);
        const int j = W.off + jl;
        const double4 pj = STAGED ? stage[jl] : b.p4[j];
        double dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
        entry_disp(e, dx, dy, dz, W, MINIMG);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 <= rc2 && r2 > (MINIMG ? 1.0e-16 : 1.0e-3))) continue;
        const int tab = ONE ? F.eam_tab[0] : F.eam_tab[ti * ns + (int)pj.w];
        if (tab < 0) continue;
        const EamTabDev &T_ = F.eamt[tab];
        const double r = sqrt(r2);
        if (r < kEamTiny || !(r < T_.cut)) continue;
        int idx;
        const double phi = eam_interp(T_.pair, 3, T_.nr, r, T_.dr, idx);
        const double dphi = eam_interp(T_.pair + 1, 3, T_.nr, r, T_.dr, idx);
        const double fs = -r * dphi;
        double fm = 0.0;
        double fpa, fpb;
        bool dens_ok;
        if (ONE) {
            fpa = fpi;
            fpb = pj.w;
            dens_ok = fpa > 0.5 * kEamNoDeriv && fpb > 0.5 * kEamNoDeriv;
        } else {
            const double rhoj = b.rho[j];
            dens_ok = rhoi >= kEamTiny && rhoj >= kEamTiny;
            fpa = dens_ok ? eam_interp(T_.emb + 1, 2, T_.nrho, rhoi, T_.drho, idx) : 0.0;
            if (dens_ok && (idx < 0 || idx >= T_.nrho)) atomicOr(b.overflow, AKMC_OV_EAM_RHO);
            fpb = dens_ok ? eam_interp(T_.emb + 1, 2, T_.nrho, rhoj, T_.drho, idx) : 0.0;
            if (dens_ok && (idx < 0 || idx >= T_.nrho)) atomicOr(b.overflow, AKMC_OV_EAM_RHO);
        }
        if (dens_ok) {
            const double drho = eam_interp(T_.pair + 2, 3, T_.nr, r, T_.dr, idx);
            fm = ((fpa + fpb) * drho) * r;
        }
        const double ft = (fs - fm) / r2;  // both orderings of the pair
        fx += ft * dx;
        fy += ft * dy;
        fz += ft * dz;
        ep += 0.5 * phi;
        if (STRESS) {
            const double h = -0.5 * ft;
            s0 += h * dx * dx;
            s1 += h * dy * dy;
            s2 += h * dz * dz;
            s3 += h * dy * dz;
            s4 += h * dx * dz;
            s5 += h * dx * dy;
        }
    }
    fx = lane_group_sum<LPA>(fx);
    fy = lane_group_sum<LPA>(fy);
    fz = lane_group_sum<LPA>(fz);
    ep = lane_group_sum<LPA>(ep);
    if (STRESS) {
        s0 = lane_group_sum<LPA>(s0);
        s1 = lane_group_sum<LPA>(s1);
        s2 = lane_group_sum<LPA>(s2);
        s3 = lane_group_sum<LPA>(s3);
        s4 = lane_group_sum<LPA>(s4);
        s5 = lane_group_sum<LPA>(s5);
    }
    if (sub != 0) return;
    b.fx[i] = fx;
    b.fy[i] = fy;
    b.fz[i] = fz;
    if (!b.pe1) energy_b<T>(b)[i] = ep;  // search micro-steps: pass 1 wrote it
    if (STRESS) {
        b.st_atom[0 * (size_t)b.A + i] = s0;
        b.st_atom[1 * (size_t)b.A + i] = s1;
        b.st_atom[2 * (size_t)b.A + i] = s2;
        b.st_atom[3 * (size_t)b.A + i] = s3;
        b.st_atom[4 * (size_t)b.A + i] = s4;
        b.st_atom[5 * (size_t)b.A + i] = s5;
    }
}

// ------------------------------------------------ Buckingham + Ewald real

// real-space Ewald and Buckingham over the full list
// (potentials/pair/Field.cpp:660-842, 1016-1201; hfunc1 :2218-2236;
// TwoBody.h:203-215).  Self-image pairs carry full weight as in the
// reference's no-image loop, other pairs half (each is listed twice).
// one real-space pair: Coulomb erfc term and Buckingham, returns the radial
// factor forc = (dE/dr)/r and the pair energies
template <typename T>
__device__ inline void ion_pair_terms(const FieldDev &F, const IonPairDev &p, double r2, double qiq, T &forc, T &ec,
                                      T &eb, int energy_path)
{
    const T tsp = T(1.1283791670955126);  // 2/sqrt(pi)
    T r2t = T(r2);
    T r = sqrt(r2t);
    T ir = T(1) / r;
    forc = T(0);
    ec = T(0);
    eb = T(0);
    if (F.do_ewald == 2) {
        // interpolate_shifted_damped_potential (pair Field.cpp:2049-2094):
        // three-point interpolation of the erfc mesh plus the shift terms
        const double rstep = 1.0 / F.dsf_del;
        const double x = double(r) * rstep;
        int point = int(x);
        if (point > F.dsf_n - 3) point = F.dsf_n - 3;
        const T dlt = T(x - double(point));
        const double *te = F.dsf_e + point, *tf = F.dsf_f + point;
        T d1 = T(te[0]) + (T(te[1]) - T(te[0])) * dlt;
        T d2 = T(te[1]) + (T(te[2]) - T(te[1])) * (dlt - T(1));
        const T omega = d1 + (d2 - d1) * dlt * T(0.5);
        d1 = T(tf[0]) + (T(tf[1]) - T(tf[0])) * dlt;
        d2 = T(tf[1]) + (T(tf[2]) - T(tf[1])) * (dlt - T(1));
        const T gamma = d1 + (d2 - d1) * dlt * T(0.5);
        const T cq = T(qiq);
        ec = cq * (omega - T(F.dsf_rfld0) + T(F.dsf_rfld1) * (r - T(F.rc)));
        forc = -cq * (gamma - T(F.dsf_rfld1) * ir);
    } else if (F.do_ewald) {
        T eta = T(F.eta);
        T cq = T(qiq);
        T xr = eta * r;
        T erfcv = erfc(xr);
        T ex = exp(-xr * xr);
        ec = cq * erfcv * ir;
        forc = cq * (-erfcv * ir * ir * ir - tsp * eta * ex * ir * ir);
    }
    if (p.on && p.form == ION_BUCK) {
        T A = sizeof(T) == 8 ? T(p.Ad) : T(p.A);
        T irho = sizeof(T) == 8 ? T(p.inv_rhod) : T(p.inv_rho);
        T Cc = sizeof(T) == 8 ? T(p.Cd) : T(p.C);
        T e1 = A * exp(-r * irho);
        T ir6 = ir * ir * ir;
        ir6 *= ir6;
        T e2 = -Cc * ir6;
        eb = e1 + e2;
        forc -= e1 * irho * ir + T(6) * e2 * ir * ir;
    } else if (p.on) {
        // the other forms with the reference's own force expressions
        // (TwoBody.h:217-276): -(dV/dr)/r for Morse and BHM, +(dV/dr)/r for
        // Lennard-Jones, and the 7-6 expression as written
        T m[6];
        for (int k = 0; k < 6; ++k) m[k] = sizeof(T) == 8 ? T(p.pd[k]) : T(p.pf[k]);
        const T ir2 = ir * ir;
        T fs = T(0);
        if (p.form == ION_MORSE) {
            const T e1 = exp(-m[1] * (r - m[2]));
            eb = m[0] * (T(1) - e1) * (T(1) - e1) - m[0];
            if (energy_path) {
                T ir12 = ir2 * ir2 * ir2;
                ir12 *= ir12;
                eb += m[3] * ir12;
            }
            fs = -(T(2) * m[0] * m[1] * (T(1) - e1) * e1) * ir;
        } else if (p.form == ION_LJONES) {
            T t6 = m[1] * ir;
            t6 = t6 * t6 * t6;
            t6 *= t6;
            const T e1 = T(4) * m[0] * t6 * t6, e2 = T(-4) * m[0] * t6;
            eb = e1 + e2;
            fs = -(T(12) * e1 + T(6) * e2) * ir2;
        } else if (p.form == ION_BHM) {
            const T ir6 = ir2 * ir2 * ir2, ir8 = ir6 * ir2;
            const T ex = exp(m[1] * (m[2] - r));
            eb = m[0] * ex - m[3] * ir6 - m[4] * ir8;
            fs = (r * m[0] * m[1] * ex - T(6) * m[3] * ir6 - T(8) * m[4] * ir8) * ir2;
        } else if (p.form == ION_SEVNSIX) {
            const T s = m[1] * ir;
            T s6 = s * s * s;
            s6 *= s6;
            eb = m[0] * s6 * s - T(2) * m[0] * s6;
            // f1 = -7 eps sigma^7 / r^8, f2 = -12 eps sigma^6 / r^7, fs = -(f1 + f2) / r^2
            fs = (T(7) * m[0] * s6 * s + T(12) * m[0] * s6) * ir * ir2;
        }
        forc -= fs;
    }
}

// real-space Ewald and Buckingham over the full list
// (potentials/pair/Field.cpp:660-842, 1016-1201; hfunc1 :2218-2236;
// TwoBody.h:203-215).  Self-image pairs carry full weight as in the
// reference's no-image loop, other pairs half (each is listed twice).
// On the FP32 path the leading near-shell entries of each list are
// evaluated in FP64: they carry the large, cancelling Coulomb and
// Buckingham terms.
// The ionic kernels give each atom one wavefront: its list (about 550
// entries at rc = 10 A) or the walker's k vectors are split over the 64
// lanes (entry or k = lane, lane + 64, ...), whose sums meet by shuffles in
// a fixed order.  256-thread workgroups of 4 atoms, dealt XCD-major, spread
// even a single 510-ion cell over 128 CUs.  The lane count is fixed, so an
// atom's sums do not depend on the batch.
constexpr int kIonLanes = 64;
constexpr int kIonAtomsPerBlock = kBlock / kIonLanes;

template <typename T, bool MINIMG, bool STRESS>
__global__ void __launch_bounds__(kBlock) k_ion_real(BatchDev b)
{
    const int nblk = (b.A + kIonAtomsPerBlock - 1) / kIonAtomsPerBlock;
    const int per = (nblk + kXCD - 1) / kXCD;
    const int blk = (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
    const int i = blk * kIonAtomsPerBlock + int(threadIdx.x) / kIonLanes, sub = int(threadIdx.x) % kIonLanes;
    if (blk >= nblk || i >= b.A) return;  // whole wavefronts leave together
    if (other_precision<T>(b, walker_of(b, i))) return;
    const WalkerDev W = b.walkers[walker_of(b, i)];
    const FieldDev &F = *b.field;
    const double xi = b.x[i], yi = b.y[i], zi = b.z[i];
    const int ti = b.type[i];
    const double rc2 = F.rc2;
    const int cnt = b.nnb[i];
    const int nnear = sizeof(T) == 8 ? cnt : b.nnear[i];
    const double qi = F.cfact * b.q[i];
    double fx = 0.0, fy = 0.0, fz = 0.0, er = 0.0, ev = 0.0;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    for (int k = sub; k < cnt; k += kIonLanes) {
        uint32_t e = nbr_entry(b, k, i);
        int jl = int(e & AKMC_NB_JMASK);
        int j = W.off + jl;
        const double4 pj = b.p4[j];
        const double qj = b.q[j];
        double dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
        entry_disp(e, dx, dy, dz, W, MINIMG);
        double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 <= rc2) || r2 < 1.0e-16) continue;
        const IonPairDev &p = F.ion[ti * F.nspec + (int)pj.w];
        double df, dec, deb;
        if (k < nnear) {
            double forc, ec, eb;
            ion_pair_terms<double>(F, p, r2, qi * qj, forc, ec, eb, b.energy_only);
            df = forc;
            dec = ec;
            deb = eb;
        } else {
            T forc, ec, eb;
            ion_pair_terms<T>(F, p, r2, qi * qj, forc, ec, eb, b.energy_only);
            df = double(forc);
            dec = double(ec);
            deb = double(eb);
        }
        double wgt = (j == i) ? 1.0 : 0.5;
        fx -= df * dx;
        fy -= df * dy;
        fz -= df * dz;
        er += wgt * dec;
        ev += wgt * deb;
        if (STRESS) {
            double h = wgt * df;
            s0 += h * dx * dx;
            s1 += h * dy * dy;
            s2 += h * dz * dz;
            s3 += h * dy * dz;
            s4 += h * dx * dz;
            s5 += h * dx * dy;
        }
    }
    fx = lane_group_sum<kIonLanes>(fx);
    fy = lane_group_sum<kIonLanes>(fy);
    fz = lane_group_sum<kIonLanes>(fz);
    er = lane_group_sum<kIonLanes>(er);
    ev = lane_group_sum<kIonLanes>(ev);
    if (STRESS) {
        s0 = lane_group_sum<kIonLanes>(s0);
        s1 = lane_group_sum<kIonLanes>(s1);
        s2 = lane_group_sum<kIonLanes>(s2);
        s3 = lane_group_sum<kIonLanes>(s3);
        s4 = lane_group_sum<kIonLanes>(s4);
        s5 = lane_group_sum<kIonLanes>(s5);
    }
    if (sub != 0) return;
    b.fx[i] = fx;
    b.fy[i] = fy;
    b.fz[i] = fz;
    b.e_a[i] = er;
    b.e_b[i] = ev;
    if (STRESS) {
        b.st_atom[0 * (size_t)b.A + i] = s0;
        b.st_atom[1 * (size_t)b.A + i] = s1;
        b.st_atom[2 * (size_t)b.A + i] = s2;
        b.st_atom[3 * (size_t)b.A + i] = s3;
        b.st_atom[4 * (size_t)b.A + i] = s4;
        b.st_atom[5 * (size_t)b.A + i] = s5;
    }
}

// ------------------------------------------------- Ewald reciprocal (FP64)

// e^{i n theta_a}, theta_a = 2 pi s_a, for n = 0..gmax by complex recurrence
__global__ void k_ewald_tables(BatchDev b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.A) return;
    if (skipped(b, walker_of(b, i))) return;
    const WalkerDev W = b.walkers[walker_of(b, i)];
    const double pos[3] = {b.x[i], b.y[i], b.z[i]};
    for (int a = 0; a < 3; ++a) {
        double th = 6.283185307179586 * pos[a] * W.invL[a];
        double s, c;
        sincos(th, &s, &c);
        double2 *t = b.ptab + ((size_t)i * 3 + a) * b.G1;
        double re = 1.0, im = 0.0;
        t[0] = make_double2(1.0, 0.0);
        for (int n = 1; n <= W.gmax; ++n) {
            double nr = re * c - im * s, ni = re * s + im * c;
            re = nr;
            im = ni;
            t[n] = make_double2(re, im);
        }
    }
}

__device__ inline double2 phase_of(const double2 *t, int G1, int nx, int ny, int nz)
{
    double2 a = t[nx];
    double2 bb = t[G1 + (ny < 0 ? -ny : ny)];
    if (ny < 0) bb.y = -bb.y;
    double2 c = t[2 * G1 + (nz < 0 ? -nz : nz)];
    if (nz < 0) c.y = -c.y;
    double2 ab = make_double2(a.x * bb.x - a.y * bb.y, a.x * bb.y + a.y * bb.x);
    return make_double2(ab.x * c.x - ab.y * c.y, ab.x * c.y + ab.y * c.x);
}

// C_k = sum_i q_i cos(k.r_i), S_k = sum_i q_i sin(k.r_i) (Ewald.cpp:596-611).
// One 1024-thread workgroup per (tile of 64 k vectors, walker): lane l of
// every wavefront owns k = 64 tile + l, wavefront v sums the atoms
// v, v + 16, ... of the walker, and the 16 wavefront partials are added in
// wavefront order through LDS (FP64 throughout; a fixed order, so the sums
// do not depend on the batch).  The per-k products g0 C and g0 S of the
// force pass are written with them.
constexpr int kSkWaves = AKMC_STAGE_BLOCK / 64;

__global__ void __launch_bounds__(AKMC_STAGE_BLOCK) k_ewald_sk(BatchDev b)
{
    __shared__ double part[2][kSkWaves][64];
    const int w = blockIdx.y;
    if (skipped(b, w)) return;
    const WalkerDev W = b.walkers[w];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    if (blockIdx.x * 64 >= W.nk) return;
    const bool has = k < W.nk;
    const int kk = W.k_off + (has ? k : 0);
    const int nx = b.kn[kk], ny = b.kn[b.nk_total + kk], nz = b.kn[2 * b.nk_total + kk];
    double C = 0.0, S = 0.0;
    if (has) {
        for (int i = W.off + wave; i < W.off + W.n; i += kSkWaves) {
            const double2 p = phase_of(b.ptab + (size_t)i * 3 * b.G1, b.G1, nx, ny, nz);
            const double q = b.q[i];
            C += q * p.x;
            S += q * p.y;
        }
    }
    part[0][wave][lane] = C;
    part[1][wave][lane] = S;
    __syncthreads();
    if (wave == 0 && has) {
        double Ct = part[0][0][lane], St = part[1][0][lane];
        for (int v = 1; v < kSkWaves; ++v) {
            Ct += part[0][v][lane];
            St += part[1][v][lane];
        }
        b.Ck[kk] = Ct;
        b.Sk[kk] = St;
    }
}

// F_i += -g0 (-q sin C + q cos S) k;  E_i = 1/2 sum_k g0 q (cos C + sin S)
// (Ewald.cpp:636-646).  One wavefront per atom, its lanes over the walker's
// k vectors (k = lane, lane + 64, ...: every load of k data is one coalesced
// 64-wide read), lane sums by shuffles.
__global__ void __launch_bounds__(kBlock) k_ewald_force(BatchDev b)
{
    const int nblk = (b.A + kIonAtomsPerBlock - 1) / kIonAtomsPerBlock;
    const int per = (nblk + kXCD - 1) / kXCD;
    const int blk = (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
    const int i = blk * kIonAtomsPerBlock + int(threadIdx.x) / kIonLanes, sub = int(threadIdx.x) % kIonLanes;
    if (blk >= nblk || i >= b.A) return;
    if (skipped(b, walker_of(b, i))) return;
    const WalkerDev W = b.walkers[walker_of(b, i)];
    const double2 *t = b.ptab + (size_t)i * 3 * b.G1;
    const double q = b.q[i];
    const int nk = b.nk_total;
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
    for (int k = sub; k < W.nk; k += kIonLanes) {
        const int kk = W.k_off + k;
        const double2 p = phase_of(t, b.G1, b.kn[kk], b.kn[nk + kk], b.kn[2 * nk + kk]);
        const double g0 = b.g0[kk];
        const double gC = g0 * b.Ck[kk], gS = g0 * b.Sk[kk];
        const double cs = q * p.x, sn = q * p.y;
        const double pf = -(-sn * gC + cs * gS);
        fx += pf * b.kv[kk];
        fy += pf * b.kv[nk + kk];
        fz += pf * b.kv[2 * nk + kk];
        e += cs * gC + sn * gS;
    }
    fx = lane_group_sum<kIonLanes>(fx);
    fy = lane_group_sum<kIonLanes>(fy);
    fz = lane_group_sum<kIonLanes>(fz);
    e = lane_group_sum<kIonLanes>(e);
    if (sub != 0) return;
    b.fx[i] += fx;
    b.fy[i] += fy;
    b.fz[i] += fz;
    b.e_c[i] = 0.5 * e;
}

// ----------------------------------------------------- walker reductions

// Workgroup sums of NV values in a fixed order: a butterfly over the 64
// lanes of each wavefront, then the NT / 64 wavefront partials in order, so
// every thread gets the same bits whatever the batch holds.
template <int NV, int NT = kBlock>
__device__ inline void block_sum(double (&v)[NV], double *sh)
{
    constexpr int kW = NT / 64;
    static_assert(NV * kW <= NT, "partials fit the scratch");
#pragma unroll
    for (int c = 0; c < NV; ++c)
        for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < NV; ++c) sh[c * kW + wave] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double t = sh[c * kW];
        for (int k = 1; k < kW; ++k) t += sh[c * kW + k];
        v[c] = t;
    }
    __syncthreads();
}

// The many-body fields' walker sums without stress (the search micro-steps):
// the two energy terms only, so a workgroup of kReduceBlock threads keeps its
// registers; fixed summation order (deterministic)
constexpr int kReduceBlock = 1024;
__global__ void __launch_bounds__(kReduceBlock) k_walker_reduce_many(BatchDev b)
{
    __shared__ double sh[kReduceBlock];
    const int w = blockIdx.x;
    if (skipped(b, w)) return;
    const WalkerDev W = b.walkers[w];
    const bool e64 = b.active[w] == 2 && b.e_a64;
    const double *ea = e64 ? b.e_a64 : b.e_a, *eb = e64 ? b.e_b64 : b.e_b;
    double v[2] = {0.0, 0.0};
    for (int i = W.off + threadIdx.x; i < W.off + W.n; i += kReduceBlock) {
        v[0] += ea[i];
        v[1] += eb[i];
    }
    block_sum<2, kReduceBlock>(v, sh);
    if (threadIdx.x == 0) {
        WalkerOut o;
        o.many = v[0];
        o.vdw = v[1];
        o.real = o.rcp = o.self = 0.0;
        o.total = o.many + o.vdw;
        for (int c = 0; c < 6; ++c) o.stress[c] = 0.0;
        b.out[w] = o;
        // the walker's full evaluation in this precision is done
        if (b.zflag) b.zflag[w] &= ~(b.active[w] == 2 ? AKMC_ZONE_FULL64 : AKMC_ZONE_FULL32);
    }
}

// one workgroup per walker; fixed summation order (deterministic)
__global__ void __launch_bounds__(kBlock) k_walker_reduce(BatchDev b, int stress)
{
    __shared__ double sh[kBlock];
    int w = blockIdx.x;
    if (skipped(b, w)) return;
    const WalkerDev W = b.walkers[w];
    const FieldDev &F = *b.field;
    // the many-body kernels keep FP64 walkers' energies apart
    const bool e64 = F.kind == 0 && b.active[w] == 2 && b.e_a64;
    const double *ea = e64 ? b.e_a64 : b.e_a, *eb = e64 ? b.e_b64 : b.e_b;
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = W.off + threadIdx.x; i < W.off + W.n; i += blockDim.x) {
        v[0] += ea[i];
        v[1] += eb[i];
        if (F.kind == 1 && F.do_ewald == 1) {
            v[2] += b.e_c[i];
            v[3] += b.q[i] * b.q[i];
        }
        if (stress)
            for (int c = 0; c < 6; ++c) v[4 + c] += b.st_atom[(size_t)c * b.A + i];
    }
    double phtot = 0.0, rs[6] = {0, 0, 0, 0, 0, 0};
    const bool rcp_sk = b.energy_only && F.kind == 1 && F.do_ewald == 1;
    if ((stress || rcp_sk) && F.kind == 1 && F.do_ewald == 1) {
        // reciprocal stress (Ewald.cpp:613-657) with the phase total from 0
        for (int k = threadIdx.x; k < W.nk; k += blockDim.x) {
            int kk = W.k_off + k;
            double C = b.Ck[kk], S = b.Sk[kk];
            double ph2 = 0.5 * (C * C + S * S);
            if (ph2 > 1.0e-16) {
                phtot += ph2 * b.g0[kk];
                double fac = 2.0 * ph2 * b.g1[kk];
                double vx = b.kv[kk], vy = b.kv[b.nk_total + kk], vz = b.kv[2 * b.nk_total + kk];
                rs[0] += fac * vx * vx;
                rs[1] += fac * vy * vy;
                rs[2] += fac * vz * vz;
                rs[3] += fac * vy * vz;
                rs[4] += fac * vx * vz;
                rs[5] += fac * vx * vy;
            }
        }
    }
    block_sum<10>(v, sh);
    double r7[7] = {phtot, rs[0], rs[1], rs[2], rs[3], rs[4], rs[5]};
    if ((stress || rcp_sk) && F.kind == 1 && F.do_ewald == 1) block_sum<7>(r7, sh);
    if (threadIdx.x == 0) {
        WalkerOut o;
        if (F.kind == 0) {
            o.many = v[0];
            o.vdw = v[1];
            o.real = o.rcp = o.self = 0.0;
            o.total = o.many + o.vdw;
        } else {
            // self term -eta/sqrt(pi) CTOEV sum q^2 (pair Field.cpp:704-720)
            o.self = F.do_ewald == 1 ? -1.1283791670955126 * 0.5 * F.cfact * F.eta * v[3] : 0.0;
            o.real = v[0] + o.self;
            o.vdw = v[1];
            // energy only: 1/2 sum_k g0 (C^2 + S^2) (Ewald::recipSpaceEnergy,
            // Ewald.cpp:368-451) instead of the per-atom sums of the force pass
            o.rcp = rcp_sk ? r7[0] : v[2];
            o.many = 0.0;
            o.total = o.rcp + o.real + o.vdw;
        }
        for (int c = 0; c < 6; ++c) o.stress[c] = stress ? v[4 + c] : 0.0;
        if (stress && F.kind == 1 && F.do_ewald == 1) {
            for (int c = 0; c < 6; ++c) o.stress[c] += r7[1 + c];
            o.stress[0] -= r7[0];
            o.stress[1] -= r7[0];
            o.stress[2] -= r7[0];
        }
        b.out[w] = o;
        // the walker's full evaluation in this precision is done
        if (b.zflag) b.zflag[w] &= ~(b.active[w] == 2 ? AKMC_ZONE_FULL64 : AKMC_ZONE_FULL32);
    }
}

inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

// ------------------------------------------------------------ launchers

static unsigned long long *g_rebuild_counter = nullptr;

extern "C" unsigned long long *akmc_rebuild_counter_ptr()
{
    if (!g_rebuild_counter) {
        if (hipMalloc(&g_rebuild_counter, sizeof(unsigned long long)) != hipSuccess) return nullptr;
        if (hipMemset(g_rebuild_counter, 0, sizeof(unsigned long long)) != hipSuccess) return nullptr;
    }
    return g_rebuild_counter;
}

static void launch_lists(const BatchDev &b, const FieldDev &hf, hipStream_t s)
{
    int gA = blocks(b.A);
    if (b.zflag) k_pack_check_zoned<<<b.W, kPackBlock, 0, s>>>(b);
    else k_pack_check<<<gA, kBlock, 0, s>>>(b);
    if (hf.min_image) {
        k_cell_bin<<<b.W, kBinBlock, 0, s>>>(b);  // records the reference positions too
        k_nbr_build_cells<<<kNbGrid, kBlock, 0, s>>>(b);
    } else {
        k_ref_positions<<<gA, kBlock, 0, s>>>(b);
        k_nbr_build_images<<<(b.A + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, s>>>(b);
    }
    // the zones of search micro-steps, with the rebuild bookkeeping
    if (b.zflag) k_zone_build<<<b.W, kBlock, 0, s>>>(b, b.rebuild_count);
    else k_rebuild_done<<<blocks(b.W), kBlock, 0, s>>>(b, b.rebuild_count);
}

// a single species pair with the exponents n = 9, m = 6 (the ONE kernels;
// other pairs take the table kernels)
static bool sc_one(const FieldDev &hf)
{
    return hf.nspec == 1 && hf.sc[0].on && hf.sc[0].n_int == 9 && hf.sc[0].m_int == 6;
}

// grid, workgroup size and dynamic LDS of a Sutton-Chen pair launch
struct ScLaunch {
    int grid, block, lds;
};

template <int LPA>
static ScLaunch sc_launch(const BatchDev &b)
{
    constexpr int li = LPA == 1 ? 0 : LPA == 2 ? 1 : LPA == 4 ? 2 : 3;
    if (b.staged) return {xcd_grid(b.nchunk_tab[li]), AKMC_STAGE_BLOCK, b.stage_bytes};
    return {xcd_grid((b.A * LPA + kBlock - 1) / kBlock), kBlock, 0};
}

// search micro-steps (minimum-image fields only) take the ZONE variants of
// pass 1, which also write the pair energies and visit active zones
template <typename T, int LPA, int U, bool ONE, bool STAGED>
static void launch_sc_density_one(const BatchDev &b, const FieldDev &hf, hipStream_t s)
{
    const ScLaunch L = sc_launch<LPA>(b);
    const int gwl = STAGED ? b.wl_grid_staged : b.wl_grid_plain;
    if (hf.min_image && b.pe1) k_sc_density<T, true, LPA, U, ONE, STAGED, true><<<gwl, L.block, L.lds, s>>>(b);
    else if (hf.min_image) k_sc_density<T, true, LPA, U, ONE, STAGED, false><<<L.grid, L.block, L.lds, s>>>(b);
    else k_sc_density<T, false, LPA, U, ONE, STAGED, false><<<L.grid, L.block, L.lds, s>>>(b);
}

template <typename T, int LPA, int U>
static void launch_sc_density(const BatchDev &b, const FieldDev &hf, hipStream_t s)
{
    if (sc_one(hf)) {
        if (b.staged) launch_sc_density_one<T, LPA, U, true, true>(b, hf, s);
        else launch_sc_density_one<T, LPA, U, true, false>(b, hf, s);
    } else {
        if (b.staged) launch_sc_density_one<T, LPA, U, false, true>(b, hf, s);
        else launch_sc_density_one<T, LPA, U, false, false>(b, hf, s);
    }
}

template <typename T, int LPA, int U, bool ONE, bool STAGED>
static void launch_sc_force_one(const BatchDev &b, const FieldDev &hf, int want_stress, hipStream_t s)
{
    const ScLaunch L = sc_launch<LPA>(b);
    const int gwl = STAGED ? b.wl_grid_staged : b.wl_grid_plain;
    // the stress variant exists for U = 1 only: U does not change results;
    // search micro-steps (minimum image, no stress) run the work list
    if (hf.min_image) {
        if (want_stress) k_sc_force<T, true, true, LPA, 1, ONE, STAGED, false><<<L.grid, L.block, L.lds, s>>>(b);
        else if (b.pe1) k_sc_force<T, true, false, LPA, U, ONE, STAGED, true><<<gwl, L.block, L.lds, s>>>(b);
        else k_sc_force<T, true, false, LPA, U, ONE, STAGED, false><<<L.grid, L.block, L.lds, s>>>(b);
    } else {
        if (want_stress) k_sc_force<T, false, true, LPA, 1, ONE, STAGED, false><<<L.grid, L.block, L.lds, s>>>(b);
        else k_sc_force<T, false, false, LPA, U, ONE, STAGED, false><<<L.grid, L.block, L.lds, s>>>(b);
    }
}

template <typename T, int LPA, int U>
static void launch_sc_force(const BatchDev &b, const FieldDev &hf, int want_stress, hipStream_t s)
{
    if (sc_one(hf)) {
        if (b.staged) launch_sc_force_one<T, LPA, U, true, true>(b, hf, want_stress, s);
        else launch_sc_force_one<T, LPA, U, true, false>(b, hf, want_stress, s);
    } else {
        if (b.staged) launch_sc_force_one<T, LPA, U, false, true>(b, hf, want_stress, s);
        else launch_sc_force_one<T, LPA, U, false, false>(b, hf, want_stress, s);
    }
}

// (lanes per atom, gathers in flight per lane) of the batch
#define AKMC_SC_UNROLL(CALL, L_)                                  \
    do {                                                          \
        if (b.unroll == 8) CALL(L_, 8);                           \
        else if (b.unroll == 4) CALL(L_, 4);                      \
        else if (b.unroll == 2) CALL(L_, 2);                      \
        else CALL(L_, 1);                                         \
    } while (0)
#define AKMC_SC_VARIANT(CALL)                                     \
    do {                                                          \
        if (b.lpa == 8) AKMC_SC_UNROLL(CALL, 8);                  \
        else if (b.lpa == 4) AKMC_SC_UNROLL(CALL, 4);             \
        else if (b.lpa == 2) AKMC_SC_UNROLL(CALL, 2);             \
        else AKMC_SC_UNROLL(CALL, 1);                             \
    } while (0)

// the EAM kernels: 4 lanes per atom
constexpr int kEamLanes = 4;

// (the staged variants read the gather records from LDS like the
// Sutton-Chen ones; the tables stay in global memory)
template <typename T, bool ONE, bool STAGED>
static void launch_eam_density_one(const BatchDev &b, const FieldDev &hf, hipStream_t s)
{
    const ScLaunch L = sc_launch<kEamLanes>(b);
    if (hf.min_image) k_eam_density<T, true, kEamLanes, ONE, STAGED><<<L.grid, L.block, L.lds, s>>>(b);
    else k_eam_density<T, false, kEamLanes, ONE, STAGED><<<L.grid, L.block, L.lds, s>>>(b);
}

template <typename T>
static void launch_eam_density(const BatchDev &b, const FieldDev &hf, hipStream_t s)
{
    if (hf.eam_one) {
        if (b.staged) launch_eam_density_one<T, true, true>(b, hf, s);
        else launch_eam_density_one<T, true, false>(b, hf, s);
    } else {
        if (b.staged) launch_eam_density_one<T, false, true>(b, hf, s);
        else launch_eam_density_one<T, false, false>(b, hf, s);
    }
}

template <typename T, bool ONE, bool STAGED>
static void launch_eam_force_one(const BatchDev &b, const FieldDev &hf, int want_stress, hipStream_t s)
{
    const ScLaunch L = sc_launch<kEamLanes>(b);
    if (hf.min_image) {
        if (want_stress) k_eam_force<T, true, true, kEamLanes, ONE, STAGED><<<L.grid, L.block, L.lds, s>>>(b);
        else k_eam_force<T, true, false, kEamLanes, ONE, STAGED><<<L.grid, L.block, L.lds, s>>>(b);
    } else {
        if (want_stress) k_eam_force<T, false, true, kEamLanes, ONE, STAGED><<<L.grid, L.block, L.lds, s>>>(b);
        else k_eam_force<T, false, false, kEamLanes, ONE, STAGED><<<L.grid, L.block, L.lds, s>>>(b);
    }
}

template <typename T>
static void launch_eam_force(const BatchDev &b, const FieldDev &hf, int want_stress, hipStream_t s)
{
    if (hf.eam_one) {
        if (b.staged) launch_eam_force_one<T, true, true>(b, hf, want_stress, s);
        else launch_eam_force_one<T, true, false>(b, hf, want_stress, s);
    } else {
        if (b.staged) launch_eam_force_one<T, false, true>(b, hf, want_stress, s);
        else launch_eam_force_one<T, false, false>(b, hf, want_stress, s);
    }
}

template <typename T>
static void launch_pass1(const BatchDev &b, const FieldDev &hf, int want_stress, hipStream_t s)
{
    int gA = xcd_grid(blocks(b.A));
    if (hf.kind == 0 && hf.eam) {
        launch_eam_density<T>(b, hf, s);
    } else if (hf.kind == 0) {
#define AKMC_DENS(L_, U_) launch_sc_density<T, L_, U_>(b, hf, s)
        AKMC_SC_VARIANT(AKMC_DENS);
#undef AKMC_DENS
    } else {
        const int gI = xcd_grid((b.A + kIonAtomsPerBlock - 1) / kIonAtomsPerBlock);
        if (hf.min_image) {
            if (want_stress) k_ion_real<T, true, true><<<gI, kBlock, 0, s>>>(b);
            else k_ion_real<T, true, false><<<gI, kBlock, 0, s>>>(b);
        } else {
            if (want_stress) k_ion_real<T, false, true><<<gI, kBlock, 0, s>>>(b);
            else k_ion_real<T, false, false><<<gI, kBlock, 0, s>>>(b);
        }
    }
}

// pass 2 with its own variant (search micro-steps)
static BatchDev pass2_view(const BatchDev &b)
{
    BatchDev v = b;
    v.lpa = b.lpa2;
    v.unroll = b.unroll2;
    v.staged = b.staged2;
    return v;
}

template <typename T>
static void launch_sc_force_variant(const BatchDev &b, const FieldDev &hf, int want_stress, hipStream_t s)
{
#define AKMC_FORCE(L_, U_) launch_sc_force<T, L_, U_>(b, hf, want_stress, s)
    AKMC_SC_VARIANT(AKMC_FORCE);
#undef AKMC_FORCE
}

template <typename T>
static void launch_pass2(const BatchDev &b, const FieldDev &hf, int want_stress, hipStream_t s)
{
    int gA = blocks(b.A);
    if (b.energy_only && hf.kind == 0 && !hf.eam) {
        if (hf.min_image) k_sc_pair_energy<T, true><<<xcd_grid(gA), kBlock, 0, s>>>(b);
        else k_sc_pair_energy<T, false><<<xcd_grid(gA), kBlock, 0, s>>>(b);
        return;
    }
    if (b.energy_only && hf.kind == 1) {
        // the reciprocal energy needs the structure factors only
        if (hf.do_ewald == 1) {
            k_ewald_tables<<<gA, kBlock, 0, s>>>(b);
            dim3 g((b.maxnk + 63) / 64, b.W);
            k_ewald_sk<<<g, AKMC_STAGE_BLOCK, 0, s>>>(b);
        }
        return;
    }
    if (hf.kind == 0 && hf.eam) {
        launch_eam_force<T>(b, hf, want_stress, s);
    } else if (hf.kind == 0) {
        launch_sc_force_variant<T>(b.lpa2 ? pass2_view(b) : b, hf, want_stress, s);
    } else if (hf.do_ewald == 1) {
        k_ewald_tables<<<gA, kBlock, 0, s>>>(b);
        dim3 g((b.maxnk + 63) / 64, b.W);
        k_ewald_sk<<<g, AKMC_STAGE_BLOCK, 0, s>>>(b);
        k_ewald_force<<<xcd_grid((b.A + kIonAtomsPerBlock - 1) / kIonAtomsPerBlock), kBlock, 0, s>>>(b);
    }
}

// The FP64 pair kernels of a mixed launch (prec 3) serve the few walkers
// under FP64 verification with their own variant: the search micro-steps'
// (lpa64, unroll64, staged64) when set, else unstaged at 4 lanes per atom
// and 4 gathers in flight (the batch's own variant is tuned for the FP32
// pass over every walker).  An FP64-only launch serves every walker and
// keeps the batch's variant and staging.
static BatchDev fp64_view(const BatchDev &b, int prec)
{
    BatchDev v = b;
    if (prec & 5) {
        v.staged = b.lpa64 ? b.staged64 : 0;
        v.lpa = b.lpa64 ? b.lpa64 : 4;  // 4 x 4 unstaged: measured best before the search variants
        v.unroll = b.lpa64 ? b.unroll64 : 4;
        v.lpa2 = 0;  // both passes
    }
    return v;
}

// one phase of a force evaluation: 0 lists, 1 pass 1 (SC density / ionic
// real space), 2 pass 2 (SC force / Ewald reciprocal), 3 walker reduction.
// prec: bit 0 launches the FP32 kernels, bit 1 the FP64 ones; each serves
// the walkers whose active mode matches.  Bit 2 marks an FP64-only launch
// that belongs to a mixed micro-step (the FP32 kernels launched apart), so
// the FP64 kernels take their mixed-step variant.
// chunk size (slots per workgroup) of a Sutton-Chen pair launch
static int chunk_atoms(const BatchDev &v) { return (v.staged ? AKMC_STAGE_BLOCK : kBlock) / v.lpa; }

// the search micro-step's work lists, for the variants the pair launches
// of this precision mask will use
static void launch_work_list(const BatchDev &b, int prec, hipStream_t s)
{
    int4 apc = make_int4(0, 0, 0, 0);
    if (prec & 1) {
        apc.x = chunk_atoms(b);
        apc.y = chunk_atoms(b.lpa2 ? pass2_view(b) : b);
    }
    if (prec & 2) {
        const BatchDev v = fp64_view(b, prec);
        apc.z = chunk_atoms(v);
        apc.w = chunk_atoms(v.lpa2 ? pass2_view(v) : v);
    }
    k_work_list<<<4, 1024, 0, s>>>(b, apc);
}

void akmc_launch_phase(const BatchDev &b, const FieldDev &hf, int prec, int want_stress, int phase,
                       hipStream_t s)
{
    switch (phase) {
    case 0:
        launch_lists(b, hf, s);
        if (b.pe1 && hf.kind == 0 && !hf.eam && hf.min_image) launch_work_list(b, prec, s);
        break;
    case 1:
        if (prec & 1) launch_pass1<float>(b, hf, want_stress, s);
        if (prec & 2) launch_pass1<double>(fp64_view(b, prec), hf, want_stress, s);
        break;
    case 2:
        if (hf.kind == 1) {
            // the reciprocal part is FP64 for every walker
            if (prec) launch_pass2<double>(b, hf, want_stress, s);
        } else {
            if (prec & 1) launch_pass2<float>(b, hf, want_stress, s);
            if (prec & 2) launch_pass2<double>(fp64_view(b, prec), hf, want_stress, s);
        }
        break;
    default:
        if (hf.kind == 0 && !want_stress) k_walker_reduce_many<<<b.W, kReduceBlock, 0, s>>>(b);
        else k_walker_reduce<<<b.W, kBlock, 0, s>>>(b, want_stress);
    }
}

void akmc_launch_forces_impl(const BatchDev &b, const FieldDev &hf, int prec, int want_stress, hipStream_t s)
{
    for (int ph = 0; ph < 4; ++ph) akmc_launch_phase(b, hf, prec, want_stress, ph, s);
}
