// Force evaluation without device allocation, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation).  No counterpart in the reference.
//
// akmc_field_forces_noalloc evaluates one Sutton-Chen configuration held in
// the caller's device arrays with the kernels of akmc_field_forces (the same
// list build, the same default kernel variant, so the same bits).  Every
// buffer those kernels need beyond the caller's inputs and outputs -- the
// walker descriptor, packed gather records, list reference positions, the
// cell grid and its work queue, the Verlet list, densities, per-atom
// energies, the walker's energy record, flags and counters -- lives in one
// fixed device array, `scratch`, of kScratchBytes (about 14 MB) per GPU,
// held until the process ends.  Every call on a GPU shares that GPU's copy
// and leaves its contents behind; each call sets what must start fresh (the
// descriptor, the rebuild, overflow and queue flags) with a small kernel of
// its own, never with hipMemsetAsync.  Calls that overlap on one GPU race on
// the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "../../include/akmc.h"
#include "akmc_internal.h"

namespace {

constexpr int kMaxAtoms = AKMC_NOALLOC_MAX_ATOMS;
constexpr int kMaxNb = AKMC_NOALLOC_MAX_NEIGHBOURS;
constexpr int kCells = AKMC_MAX_WALKER_CELLS;
constexpr int kBlock = 256;

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// byte offsets of the parts of scratch
constexpr size_t kWalker = 0;                                          // WalkerDev
constexpr size_t kOut = align256(kWalker + sizeof(WalkerDev));         // WalkerOut
constexpr size_t kFlags = align256(kOut + sizeof(WalkerOut));          // rebuild, active, overflow, queue length
constexpr size_t kCount = align256(kFlags + 4 * sizeof(int));          // rebuild counter
constexpr size_t kWk = align256(kCount + sizeof(unsigned long long));  // atom -> walker (all 0)
constexpr size_t kP4 = align256(kWk + kMaxAtoms * sizeof(int));
constexpr size_t kXr = align256(kP4 + kMaxAtoms * sizeof(double4));
constexpr size_t kNnb = align256(kXr + 3 * kMaxAtoms * sizeof(double));
constexpr size_t kNnear = align256(kNnb + kMaxAtoms * sizeof(int));
constexpr size_t kCellAtoms = align256(kNnear + kMaxAtoms * sizeof(int));
constexpr size_t kRho = align256(kCellAtoms + kMaxAtoms * sizeof(int));
constexpr size_t kIsr = align256(kRho + kMaxAtoms * sizeof(double));
constexpr size_t kIsr64 = align256(kIsr + kMaxAtoms * sizeof(float));
constexpr size_t kEa = align256(kIsr64 + kMaxAtoms * sizeof(double));
constexpr size_t kEb = align256(kEa + kMaxAtoms * sizeof(double));
constexpr size_t kEc = align256(kEb + kMaxAtoms * sizeof(double));
constexpr size_t kCellCount = align256(kEc + kMaxAtoms * sizeof(double));
constexpr size_t kCellStart = align256(kCellCount + kCells * sizeof(int));
constexpr size_t kCellWk = align256(kCellStart + kCells * sizeof(int));
constexpr size_t kCellWork = align256(kCellWk + kCells * sizeof(int));
constexpr size_t kNbr = align256(kCellWork + (kCells + kMaxAtoms / 16 + 1) * sizeof(int));
constexpr size_t kScratchBytes =
    align256(kNbr + (size_t)(kMaxNb + AKMC_NB_PAD_ROWS) * kMaxAtoms * sizeof(uint32_t));

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// what must start fresh each call: the walker descriptor (a kernel argument,
// so a captured launch carries it by value), a list rebuild, no overflow,
// an empty work queue, every atom and cell in walker 0
__global__ void k_noalloc_init(BatchDev d, WalkerDev wd, int mode, int ncells)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *const_cast<WalkerDev *>(d.walkers) = wd;
        d.rebuild[0] = 1;
        d.active[0] = mode;
        *d.overflow = 0;
        *d.cell_work_n = 0;
    }
    if (i < d.A) const_cast<int *>(d.wk)[i] = 0;
    if (i < ncells) const_cast<int *>(d.cell_wk)[i] = 0;
}

// the walker's energy record into the caller's akmc_energy, the overflow
// word into the caller's flags
__global__ void k_noalloc_finish(BatchDev d, akmc_energy *e, int *flags)
{
    if (threadIdx.x != 0) return;
    if (e) {
        const WalkerOut o = d.out[0];
        e->total = o.total;
        e->rcp = o.rcp;
        e->real = o.real;
        e->many = o.many;
        e->vdw = o.vdw;
        e->ext = 0.0;
    }
    if (flags) *flags = *d.overflow;
}

std::mutex g_mu;
constexpr int kMaxDevices = 64;
unsigned char *g_scratch[kMaxDevices] = {};

// the device address of this GPU's scratch array, looked up once per device
int scratch_of_current(unsigned char **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return akmc_fail(AKMC_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return akmc_fail(AKMC_EUNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return akmc_fail(AKMC_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[dev];
    return AKMC_SUCCESS;
}

template <typename T>
T *at(unsigned char *base, size_t off)
{
    return reinterpret_cast<T *>(base + off);
}

}  // namespace

extern "C" unsigned long akmc_noalloc_scratch_bytes(void) { return (unsigned long)kScratchBytes; }

extern "C" int akmc_field_forces_noalloc(akmc_field_t f, const double *x, const double *y, const double *z,
                                         const int *type, int n, const double lat[9], double *fx, double *fy,
                                         double *fz, akmc_energy *energy, int *flags, void *stream)
{
    if (!f || !x || !y || !z || !type || !fx || !fy || !fz || !lat)
        return akmc_fail(AKMC_EINVAL, "bad arguments");
    if (n <= 0 || n > kMaxAtoms)
        return akmc_fail(AKMC_ECAPACITY, "akmc_field_forces_noalloc takes 1 to %d atoms, got %d", kMaxAtoms, n);
    FieldView v{};
    int rc = akmc_field_view(f, &v);
    if (rc != AKMC_SUCCESS) return rc;
    if (v.kind != 0 || v.eam)
        return akmc_fail(AKMC_EUNSUPPORTED, "akmc_field_forces_noalloc serves Sutton-Chen fields only");
    if (!v.dev) return akmc_fail(AKMC_EINVAL, "call akmc_field_setup before akmc_field_forces_noalloc");
    for (int a = 0; a < 9; ++a)
        if (a % 4 != 0 && lat[a] != 0.0) return akmc_fail(AKMC_EUNSUPPORTED, "orthorhombic cells only");
    // the walker descriptor and cell grid of a one-walker batch (as batch_build)
    WalkerDev wd{};
    const double rl = v.rc + v.skin;
    wd.off = 0;
    wd.n = n;
    wd.cell_off = 0;
    const double L[3] = {lat[0], lat[4], lat[8]};
    for (int a = 0; a < 3; ++a) {
        if (!(L[a] > 0.0)) return akmc_fail(AKMC_EINVAL, "cell lengths must be positive");
        if (v.min_image && v.rc > 0.5 * L[a])
            return akmc_fail(AKMC_EINVAL, "cutoff %g exceeds half the cell (%g)", v.rc, L[a]);
        wd.L[a] = L[a];
        wd.invL[a] = 1.0 / L[a];
        wd.nc[a] = std::min(AKMC_MAX_CELLS_AXIS, std::max(1, (int)std::floor(L[a] / rl)));
        wd.shells[a] = (int)(rl / L[a]) + 1;
        if (wd.shells[a] > 7) return akmc_fail(AKMC_ECAPACITY, "more than 7 image shells");
    }
    const int ncells = wd.nc[0] * wd.nc[1] * wd.nc[2];
    unsigned char *s = nullptr;
    rc = scratch_of_current(&s);
    if (rc != AKMC_SUCCESS) return rc;

    BatchDev d{};
    d.A = n;
    d.W = 1;
    d.maxnb = kMaxNb;
    d.ncells_total = ncells;
    d.G1 = 1;
    d.lpa = 1;  // the default variant of a fresh batch, as akmc_field_forces runs
    d.unroll = 1;
    d.staged = 0;
    d.nb16 = (v.min_image && n <= 65535) ? 1 : 0;
    d.field = v.dev;
    d.walkers = at<WalkerDev>(s, kWalker);
    d.wk = at<int>(s, kWk);
    d.x = const_cast<double *>(x);  // read only by the force kernels
    d.y = const_cast<double *>(y);
    d.z = const_cast<double *>(z);
    d.p4 = at<double4>(s, kP4);
    d.xr = at<double>(s, kXr);
    d.yr = d.xr + kMaxAtoms;
    d.zr = d.yr + kMaxAtoms;
    d.type = type;
    d.fx = fx;
    d.fy = fy;
    d.fz = fz;
    d.nbr = at<uint32_t>(s, kNbr);
    d.nnb = at<int>(s, kNnb);
    d.nnear = at<int>(s, kNnear);
    int *fl = at<int>(s, kFlags);
    d.rebuild = fl;
    d.active = fl + 1;
    d.overflow = fl + 2;
    d.cell_work_n = fl + 3;
    d.rebuild_count = at<unsigned long long>(s, kCount);
    d.cell_count = at<int>(s, kCellCount);
    d.cell_start = at<int>(s, kCellStart);
    d.cell_atoms = at<int>(s, kCellAtoms);
    d.cell_wk = at<int>(s, kCellWk);
    d.cell_work = at<int>(s, kCellWork);
    d.rho = at<double>(s, kRho);
    d.inv_sqrt_rho = at<float>(s, kIsr);
    d.isr64 = at<double>(s, kIsr64);
    d.e_a = at<double>(s, kEa);
    d.e_b = at<double>(s, kEb);
    d.e_c = at<double>(s, kEc);
    d.out = at<WalkerOut>(s, kOut);
    const int mode = v.fp64 ? 2 : 1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int grid = (std::max(n, ncells) + kBlock - 1) / kBlock;
    k_noalloc_init<<<grid, kBlock, 0, st>>>(d, wd, mode, ncells);
    akmc_launch_forces_impl(d, *v.host, mode, 0, st);
    k_noalloc_finish<<<1, 64, 0, st>>>(d, energy, flags);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return akmc_fail(AKMC_EHIP, "akmc_field_forces_noalloc launch: %s", hipGetErrorString(e));
    return AKMC_SUCCESS;
}
