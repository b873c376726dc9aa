// Walker state machines for gfx950: FIRE relax (optim/Relax.cpp:44-315),
// the dimer saddle search (rotation + FIRE translation), ART
// (optim/ART.cpp) and the FP64 verification of converged saddles, advanced
// by one logic kernel after every batched force evaluation.
//
// One workgroup owns one walker.  Every vector operation of the walker is a
// strided loop of its 256 threads over the walker's atoms (relax) or free
// atoms (saddle searches) followed by a fixed-order reduction (wavefront
// butterfly, then the four wavefront partials in order), so a walker's
// trajectory is bitwise the same whether it runs alone or in a batch of
// hundreds.  Vectors are FP64 [3][A] arrays (component c of atom i at
// c*A + i).
#include "akmc_internal.h"

#include <math.h>

namespace {

// workgroup size of the walker state machines: 256 threads, or 1024 in the
// wide build of this file (search_kernels_wide.hip) for batches of walkers
// so large that the strided loops of 256 threads dominate a micro-step
#ifndef AKMC_LOGIC_BLOCK
#define AKMC_LOGIC_BLOCK 256
#endif
constexpr int kBlock = AKMC_LOGIC_BLOCK;

// Block-wide sums and maxima in a fixed order: a butterfly over the 64
// lanes of each wavefront, then the four wavefront partials in order, so
// every thread gets the same bits and a walker's result does not depend on
// the rest of the batch.
constexpr int kWaves = kBlock / 64;

template <int NV>
__device__ inline void bsum(double (&v)[NV], double *sh)
{
    for (int c = 0; c < NV; ++c)
        for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < NV; ++c) sh[c * kWaves + wave] = v[c];
    __syncthreads();
    for (int c = 0; c < NV; ++c) {
        double t = sh[c * kWaves];
        for (int k = 1; k < kWaves; ++k) t += sh[c * kWaves + k];
        v[c] = t;
    }
    __syncthreads();
}

__device__ inline double bmax(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[wave] = v;
    __syncthreads();
    double t = sh[0];
    for (int k = 1; k < kWaves; ++k) t = fmax(t, sh[k]);
    __syncthreads();
    return t;
}

struct Ctx {
    const BatchDev &b;
    const SearchDev &s;
    int off, n, A;
    int nact;  // free atoms of the walker: s.act[off .. off + nact)
    // the walker has frozen atoms: its relax loops visit the free atoms and
    // carry the frozen atoms' common velocity as WalkerState.vf
    __device__ bool part() const { return nact < n; }
    __device__ int nfrozen() const { return n - nact; }
    __device__ double f(int c, int i) const
    {
        if (s.frozen[i]) return 0.0;
        return c == 0 ? b.fx[i] : (c == 1 ? b.fy[i] : b.fz[i]);
    }
    __device__ void probe(int i, double px, double py, double pz) const
    {
        b.x[i] = px;
        b.y[i] = py;
        b.z[i] = pz;
    }
};

// every atom of the walker (FIRE relax keeps the reference's all-atom sums)
#define FOR_ATOMS(i) for (int i = C_.off + (int)threadIdx.x; i < C_.off + C_.n; i += kBlock)
// the walker's free atoms only: the vectors of the saddle searches span the
// active region (optim/SaddleSearch.cpp vectorSize = 3 N_active); frozen
// components are zero and frozen probe positions stay at X
#define FOR_ACTIVE(i)                                                                   \
    for (int t_##i = (int)threadIdx.x, i = 0;                                           \
         t_##i < C_.nact && ((i = C_.s.act[C_.off + t_##i]), true); t_##i += kBlock)
// the per-iteration relax loops: every atom, or the free atoms when the
// walker has frozen atoms.  A frozen atom has no force and never moves; its
// velocity, which enters the centre-of-mass correction and |V|^2 of the
// reference's all-atom sums (Relax.cpp:125-160, :964-996), is the same for
// every frozen atom, so it is kept once per walker.
#define FOR_RELAX(i)                                                                          \
    for (int t_##i = (int)threadIdx.x, i = 0;                                                 \
         t_##i < (C_.part() ? C_.nact : C_.n) &&                                              \
         ((i = C_.part() ? C_.s.act[C_.off + t_##i] : C_.off + t_##i), true);                 \
         t_##i += kBlock)
#define V3(arr, c, i) (arr)[(size_t)(c) * C_.A + (i)]

// probe = X + a * D (D a [3][A] vector) or X when D is null
__device__ inline void write_probe(const Ctx &C_, const double *D, double a)
{
    FOR_ACTIVE(i)
    {
        double px = V3(C_.s.X, 0, i), py = V3(C_.s.X, 1, i), pz = V3(C_.s.X, 2, i);
        if (D) {
            px += a * V3(D, 0, i);
            py += a * V3(D, 1, i);
            pz += a * V3(D, 2, i);
        }
        C_.probe(i, px, py, pz);
    }
}

// a walker that ends stops being evaluated by the force kernels
__device__ inline void finish(const Ctx &C_, WalkerState &w)
{
    w.state = ST_DONE;
    if (threadIdx.x == 0) C_.b.active[blockIdx.x] = 0;
    write_probe(C_, nullptr, 0.0);
}

// ------------------------------------------------------------------ basin 2

__device__ inline double wrap(double d, double L, double iL) { return d - L * rint(d * iL); }

// the kick of checkForNewBasin (kmc/KineticMonteCarlo.cpp:1646-1683): with
// Delta the minimum-image displacement of every atom from the basin to the
// saddle, free atoms move from the saddle by norm * Delta, where
// norm = basin2Delta * |Delta|; the saddle is kept in Xs and the walker
// FIRE-relaxes from there with the search's frozen mask
// (KineticMonteCarlo.cpp:1686)
__device__ __attribute__((always_inline)) void enter_basin2(const Ctx &C_, WalkerState &w, double *sh)
{
    const WalkerDev &W = C_.b.walkers[blockIdx.x];
    double acc[1] = {0.0};
    FOR_ATOMS(i)
    {
        for (int c = 0; c < 3; ++c) {
            const double x = V3(C_.s.X, c, i);
            const double d = wrap(x - V3(C_.s.basin, c, i), W.L[c], W.invL[c]);
            V3(C_.s.Xs, c, i) = x;
            V3(C_.s.T, c, i) = d;
            acc[0] += d * d;
        }
    }
    bsum<1>(acc, sh);
    const double norm = sqrt(acc[0]) * C_.s.basin2_delta;
    FOR_ATOMS(i)
    {
        const bool fr = C_.s.frozen[i] != 0;
        for (int c = 0; c < 3; ++c) {
            if (!fr) V3(C_.s.X, c, i) += norm * V3(C_.s.T, c, i);
            V3(C_.s.V, c, i) = 0.0;
        }
    }
    w.kind = WK_BASIN2;
    w.state = ST_R_INIT;
    w.it = 0;
    w.nstep = 1;  // Relax.cpp:49
    w.nuphill = 0;
    w.vf[0] = w.vf[1] = w.vf[2] = 0.0;
    w.dt = C_.s.rp.initTimeStep;
    w.alpha = C_.s.rp.alphaStart;
    w.b2_calls = 0;
    if (threadIdx.x == 0) C_.b.active[blockIdx.x] = C_.s.default_mode;
    __syncthreads();
    write_probe(C_, nullptr, 0.0);
}

// a converged search that is not verified (or whose verification ended)
// goes on to its basin 2 when that stage is on
__device__ __attribute__((always_inline)) void end_converged(const Ctx &C_, WalkerState &w, double *sh)
{
    if (C_.s.basin2_on && C_.s.basin) enter_basin2(C_, w, sh);
    else finish(C_, w);
}

// a converged saddle is re-verified in FP64 when verification is enabled
__device__ __attribute__((always_inline)) void enter_verify_or_finish(const Ctx &C_, WalkerState &w, double curvature, double *sh)
{
    w.e32 = w.E0;
    w.c32 = curvature;
    if (C_.s.verify_nl > 0) {
        w.kind = WK_VERIFY;
        w.state = ST_V_F0;
        w.lk = 0;
        if (threadIdx.x == 0) C_.b.active[blockIdx.x] = 2;
        write_probe(C_, nullptr, 0.0);
    } else {
        end_converged(C_, w, sh);
    }
}

// ------------------------------------------------------------------ relax

__device__ __attribute__((always_inline)) void relax_end(const Ctx &C_, WalkerState &w, double *sh, int status);

// the part of a FIRE / FIRE2 iteration before its force call: the power
// test and the dt / alpha schedule, the velocity mixing and the first half
// of the integrator.  FIRE: optim/Relax.cpp:125-225; FIRE2: :429-560, which
// mixes inside the integration loop with the alpha of before the schedule
// (the same values), counts uphill steps (failure past N_maxUphill),
// changes dt and alpha on an uphill step only from iteration N_delaySteps
// on when initialDelay is set (then clamped at minTimeStep), and puts the
// free atoms back by half a step of the new dt before zeroing the
// velocities.  Semi-implicit Euler by default; velocity Verlet
// (verletIntegration, :194-225 / :519-560) kicks the free atoms' velocities
// by half a step of the old forces and zeroes the frozen ones before the
// drift.
__device__ __attribute__((always_inline)) void relax_pre(const Ctx &C_, WalkerState &w, double *sh)
{
    const RelaxParamsDev &p = C_.s.rp;
    const bool fire2 = p.method == 1;
    double acc[3] = {0, 0, 0};  // P, |F|^2, |V|^2
    FOR_RELAX(i)
    {
        for (int c = 0; c < 3; ++c) {
            double F = V3(C_.s.F0, c, i), v = V3(C_.s.V, c, i);
            acc[0] += F * v;
            acc[1] += F * F;
            acc[2] += v * v;
        }
    }
    bsum<3>(acc, sh);
    if (C_.part()) acc[2] += C_.nfrozen() * (w.vf[0] * w.vf[0] + w.vf[1] * w.vf[1] + w.vf[2] * w.vf[2]);
    double P = acc[0];
    double af = w.alpha * sqrt(acc[2]) / sqrt(acc[1]);
    double oma = 1.0 - w.alpha;
    bool down = P > 0;
    if (down) {
        if (w.nstep > p.N_min) {
            w.dt = fmin(w.dt * p.stepIncrease, p.maxTimeStep);
            w.alpha *= p.alphaDecrease;
        }
        w.nstep += 1;
        w.nuphill = 0;
    } else if (fire2) {
        w.nstep = 0;
        w.nuphill += 1;
        if (w.nuphill > p.N_maxUphill) {  // Relax.cpp:446-459: positions as they are
            relax_end(C_, w, sh, 1);
            return;
        }
        if (!p.initialDelay) {
            w.dt *= p.stepDecrease;
            w.alpha = p.alphaStart;
        } else if (w.it >= p.N_delaySteps) {
            w.dt *= p.stepDecrease;
            if (w.dt < p.minTimeStep) w.dt = p.minTimeStep;
            w.alpha = p.alphaStart;
        }
    } else {
        w.nstep = 0;
        w.dt *= p.stepDecrease;
        w.alpha = p.alphaStart;
    }
    const double dt = w.dt, fact = dt * p.timeStepUnit, half = 0.5 * dt * p.timeStepUnit;
    // frozen velocities: kept when mixing, zeroed on an uphill step and by Verlet
    if (!down || p.verlet) w.vf[0] = w.vf[1] = w.vf[2] = 0.0;
    FOR_RELAX(i)
    {
        bool fr = C_.s.frozen[i] != 0;
        for (int c = 0; c < 3; ++c) {
            double v = V3(C_.s.V, c, i), F = V3(C_.s.F0, c, i);
            if (down) {
                if (!fr) v = oma * v + af * F;
            } else {
                if (fire2 && !fr) V3(C_.s.X, c, i) -= 0.5 * dt * v;
                v = 0.0;
            }
            if (p.verlet) {
                if (!fr) {
                    v += half * F;
                    V3(C_.s.X, c, i) += dt * v;
                } else {
                    v = 0.0;
                }
            } else if (!fr) {
                V3(C_.s.X, c, i) += dt * v;
                v += fact * F;
            }
            V3(C_.s.V, c, i) = v;
        }
    }
    __syncthreads();
    write_probe(C_, nullptr, 0.0);
}

// the end of a FIRE relax.  A basin-2 relax is then compared with the
// basin atom by atom (Basis::checkBasin, Common/Basis.cpp:731-781): a new
// basin when some atom moved by at least basinRadius and none by more than
// maxRadius, i.e. when the largest minimum-image displacement lies in
// [basinRadius, maxRadius].  Its relax status is kept apart from the
// search's, and, as KineticMonteCarlo.cpp:1692-1710 does, does not enter
// the decision.  With final64 on, one FP64 energy evaluation at the relaxed
// positions follows (their forces were just evaluated, so the probes
// already sit at X).
__device__ __attribute__((always_inline)) void relax_end(const Ctx &C_, WalkerState &w, double *sh, int status)
{
    if (w.kind == WK_BASIN2) {
        const WalkerDev &W = C_.b.walkers[blockIdx.x];
        double m = 0.0;
        FOR_ATOMS(i)
        {
            double r2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double d = wrap(V3(C_.s.basin, c, i) - V3(C_.s.X, c, i), W.L[c], W.invL[c]);
                r2 += d * d;
            }
            m = fmax(m, r2);
        }
        m = bmax(m, sh);
        w.b2_status = status;
        w.b2_iters = w.it;
        w.eb2 = w.E_last;
        w.b2_maxd2 = m;
        w.b2_new = (m >= C_.s.basin_r2 && m <= C_.s.max_r2) ? 1 : 0;
    } else {
        w.status = status;
    }
    if (C_.s.final64) {
        w.state = ST_R_E64;
        if (threadIdx.x == 0) C_.b.active[blockIdx.x] = 2;
        return;
    }
    finish(C_, w);
}

__device__ __attribute__((always_inline)) void relax_logic(const Ctx &C_, WalkerState &w, double *sh)
{
    const RelaxParamsDev &p = C_.s.rp;
    if (w.state == ST_R_E64) {
        const double e = C_.b.out[blockIdx.x].total;
        if (w.kind == WK_BASIN2) w.eb2_64 = e;
        else w.e64 = e;
        finish(C_, w);
        return;
    }
    // forces at X (frozen atoms zeroed, Relax.cpp:230; with frozen atoms only
    // the free atoms' forces are read)
    FOR_RELAX(i)
    {
        for (int c = 0; c < 3; ++c) V3(C_.s.F0, c, i) = C_.f(c, i);
    }
    if (w.kind == WK_BASIN2) w.b2_calls += 1;
    else w.fcalls += 1;
    w.E_last = C_.b.out[blockIdx.x].total;
    if (w.kind != WK_BASIN2) w.E0 = w.E_last;  // a basin-2 relax keeps the saddle energy in E0
    __syncthreads();
    if (w.state == ST_R_INIT) {
        if (p.method == 2) {  // calculateConfig only (Relax.cpp:33-39)
            relax_end(C_, w, sh, 0);
            return;
        }
        w.it = 1;
        w.state = ST_R_STEP;
        relax_pre(C_, w, sh);
        return;
    }
    if (p.verlet) {
        // second half kick with the new forces; frozen velocities stay zero
        // (Relax.cpp:233-257, :569-585)
        const double half = 0.5 * w.dt * p.timeStepUnit;
        FOR_RELAX(i)
        {
            const bool fr = C_.s.frozen[i] != 0;
            for (int c = 0; c < 3; ++c) {
                if (!fr) V3(C_.s.V, c, i) += half * V3(C_.s.F0, c, i);
                else V3(C_.s.V, c, i) = 0.0;
            }
        }
        w.vf[0] = w.vf[1] = w.vf[2] = 0.0;
        __syncthreads();
    }
    // centre-of-mass velocity removal over all atoms (Relax.cpp:964-996)
    double m[3] = {0, 0, 0};
    FOR_RELAX(i)
    {
        for (int c = 0; c < 3; ++c) m[c] += V3(C_.s.V, c, i);
    }
    bsum<3>(m, sh);
    if (C_.part()) {
        for (int c = 0; c < 3; ++c) {
            m[c] += C_.nfrozen() * w.vf[c];
            w.vf[c] -= m[c] / (double)C_.n;
        }
    }
    double fn2 = 0.0, fm2 = 0.0;
    FOR_RELAX(i)
    {
        double q = 0.0;
        for (int c = 0; c < 3; ++c) {
            V3(C_.s.V, c, i) -= m[c] / (double)C_.n;
            double F = V3(C_.s.F0, c, i);
            q += F * F;
        }
        fn2 += q;
        fm2 = fmax(fm2, q);
    }
    double t[1] = {fn2};
    bsum<1>(t, sh);
    fm2 = bmax(fm2, sh);
    if (sqrt(t[0]) < p.normTol || sqrt(fm2) < p.forceTol) {
        relax_end(C_, w, sh, 0);
        return;
    }
    if (w.it >= p.maxIter) {
        relax_end(C_, w, sh, 1);
        return;
    }
    w.it += 1;
    relax_pre(C_, w, sh);
}

// ------------------------------------------------------------------ dimer

__device__ __attribute__((always_inline)) void dimer_translate(const Ctx &C_, WalkerState &w, double *sh);

// rotation bookkeeping: decide between another trial rotation and the
// translation (the rotation loop of the CPU oracle's dimer)
__device__ __attribute__((always_inline)) void dimer_rot_eval(const Ctx &C_, WalkerState &w, double *sh)
{
    const DimerParamsDev &p = C_.s.dp;
    const double D = p.dimerDisplacement;
    double acc[1] = {0.0};
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) acc[0] += (V3(C_.s.F0, c, i) - V3(C_.s.F1, c, i)) / D * V3(C_.s.N, c, i);
    }
    bsum<1>(acc, sh);
    double Cv = acc[0];
    double g2[1] = {0.0};
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            double gp = (V3(C_.s.F0, c, i) - V3(C_.s.F1, c, i)) / D - Cv * V3(C_.s.N, c, i);
            g2[0] += gp * gp;
        }
    }
    bsum<1>(g2, sh);
    double g = sqrt(g2[0]);
    w.C = Cv;
    w.g = g;
    if (g < p.normTol || w.nrot >= p.maxRotations) {
        dimer_translate(C_, w, sh);
        return;
    }
    double t1 = 0.5 * atan(g / fabs(Cv));
    if (t1 < p.deltaTheta) {
        dimer_translate(C_, w, sh);
        return;
    }
    w.t1 = t1;
    double ct = cos(t1), st = sin(t1);
    FOR_ACTIVE(i)
    {
        double pr[3];
        for (int c = 0; c < 3; ++c) {
            double gp = (V3(C_.s.F0, c, i) - V3(C_.s.F1, c, i)) / D - Cv * V3(C_.s.N, c, i);
            double Tc = -gp / g;
            V3(C_.s.T, c, i) = Tc;
            pr[c] = V3(C_.s.X, c, i) + D * (V3(C_.s.N, c, i) * ct + Tc * st);
        }
        C_.probe(i, pr[0], pr[1], pr[2]);
    }
    w.state = ST_D_F1T;
}

// translation by FIRE on the effective force; convergence test
__device__ __attribute__((always_inline)) void dimer_translate(const Ctx &C_, WalkerState &w, double *sh)
{
    const DimerParamsDev &p = C_.s.dp;
    const double D = p.dimerDisplacement;
    double acc[2] = {0.0, 0.0};  // C*D, F0.n
    double fm2 = 0.0;
    FOR_ACTIVE(i)
    {
        double q = 0.0;
        for (int c = 0; c < 3; ++c) {
            double F0 = V3(C_.s.F0, c, i), Nn = V3(C_.s.N, c, i);
            acc[0] += (F0 - V3(C_.s.F1, c, i)) * Nn;
            acc[1] += F0 * Nn;
            q += F0 * F0;
        }
        fm2 = fmax(fm2, q);
    }
    bsum<2>(acc, sh);
    fm2 = bmax(fm2, sh);
    double Cv = acc[0] / D, fn = acc[1];
    w.C = Cv;
    if (sqrt(fm2) < p.forceTol && Cv < p.maxEigenvalue) {
        w.status = 0;
        enter_verify_or_finish(C_, w, Cv, sh);
        return;
    }
    const bool convex = Cv > 0.0;
    double a3[3] = {0, 0, 0};  // Fe.V, |V|^2, |Fe|^2
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            double Nn = V3(C_.s.N, c, i);
            double Fe = convex ? -fn * Nn : V3(C_.s.F0, c, i) - 2.0 * fn * Nn;
            double v = V3(C_.s.V, c, i);
            a3[0] += Fe * v;
            a3[1] += v * v;
            a3[2] += Fe * Fe;
        }
    }
    bsum<3>(a3, sh);
    bool down = a3[0] > 0.0;
    double vn = sqrt(a3[1]), fe = sqrt(a3[2]), al = w.alpha;
    if (down) {
        if (w.nstep > p.N_min) {
            w.dt = fmin(w.dt * p.stepIncrease, p.maxTimeStep);
            w.alpha *= p.alphaDecrease;
        }
        w.nstep += 1;
    } else {
        w.dt *= p.stepDecrease;
        w.alpha = p.alphaStart;
        w.nstep = 0;
    }
    const double dt = w.dt, fact = dt * p.timeStepUnit;
    double dm2 = 0.0;
    FOR_ACTIVE(i)
    {
        double q = 0.0;
        for (int c = 0; c < 3; ++c) {
            double Nn = V3(C_.s.N, c, i);
            double Fe = convex ? -fn * Nn : V3(C_.s.F0, c, i) - 2.0 * fn * Nn;
            double v = down ? (1.0 - al) * V3(C_.s.V, c, i) + al * vn * Fe / fe : 0.0;
            v += fact * Fe;
            V3(C_.s.V, c, i) = v;
            q += (dt * v) * (dt * v);
        }
        dm2 = fmax(dm2, q);
    }
    dm2 = bmax(dm2, sh);
    double dmax = sqrt(dm2);
    double sc = dmax > p.maxStep ? p.maxStep / dmax : 1.0;
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            double v = V3(C_.s.V, c, i) * sc;
            V3(C_.s.V, c, i) = v;
            V3(C_.s.X, c, i) += dt * v;
        }
    }
    w.it += 1;
    if (w.it > p.maxParIter) {
        w.status = 1;
        __syncthreads();
        finish(C_, w);
        return;
    }
    w.state = ST_D_F0;
    __syncthreads();
    write_probe(C_, nullptr, 0.0);
}

__device__ __attribute__((always_inline)) void dimer_logic(const Ctx &C_, WalkerState &w, double *sh)
{
    const DimerParamsDev &p = C_.s.dp;
    const double D = p.dimerDisplacement;
    w.fcalls += 1;
    if (w.state == ST_D_F0) {
        FOR_ACTIVE(i)
        {
            for (int c = 0; c < 3; ++c) V3(C_.s.F0, c, i) = C_.f(c, i);
        }
        w.E0 = C_.b.out[blockIdx.x].total;
        w.state = ST_D_F1;
        __syncthreads();
        write_probe(C_, C_.s.N, D);
        return;
    }
    if (w.state == ST_D_F1) {
        FOR_ACTIVE(i)
        {
            for (int c = 0; c < 3; ++c) V3(C_.s.F1, c, i) = C_.f(c, i);
        }
        w.nrot = 0;
        __syncthreads();
        dimer_rot_eval(C_, w, sh);
        return;
    }
    // ST_D_F1T: harmonic fit of the curvature over the trial angle and
    // extrapolation of the end force to the optimum angle
    const double t1 = w.t1, ct1 = cos(t1), st1 = sin(t1);
    double acc[1] = {0.0};
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            double nt = V3(C_.s.N, c, i) * ct1 + V3(C_.s.T, c, i) * st1;
            acc[0] += (V3(C_.s.F0, c, i) - C_.f(c, i)) * nt;
        }
    }
    bsum<1>(acc, sh);
    double Ct = acc[0] / D;
    double b1 = -w.g;
    double a1 = (w.C - Ct + b1 * sin(2.0 * t1)) / (1.0 - cos(2.0 * t1));
    double th = 0.5 * atan(b1 / a1);
    const double HALFPI = 1.5707963267948966;
    if (a1 * cos(2.0 * th) + b1 * sin(2.0 * th) > 0.0) th += HALFPI;
    if (th > HALFPI) th -= 2.0 * HALFPI;
    double s0 = sin(t1 - th) / st1, s1 = sin(th) / st1;
    double s2 = 1.0 - cos(th) - sin(th) * tan(0.5 * t1);
    double cth = cos(th), sth = sin(th);
    double nn[1] = {0.0};
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            double F1 = V3(C_.s.F1, c, i) * s0 + C_.f(c, i) * s1 + V3(C_.s.F0, c, i) * s2;
            V3(C_.s.F1, c, i) = F1;
            double Nn = V3(C_.s.N, c, i) * cth + V3(C_.s.T, c, i) * sth;
            V3(C_.s.N, c, i) = Nn;
            nn[0] += Nn * Nn;
        }
    }
    bsum<1>(nn, sh);
    double inv = 1.0 / sqrt(nn[0]);
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) V3(C_.s.N, c, i) *= inv;
    }
    w.nrot += 1;
    __syncthreads();
    dimer_rot_eval(C_, w, sh);
}


// ------------------------------------------------------------ verification

// number of eigenvalues of the symmetric tridiagonal (d, e) below x (Sturm
// sequence of the LDL^T pivots)
__device__ inline int sturm_count(int n, const double *d, const double *e, double x)
{
    int c = 0;
    double q = d[0] - x;
    for (int i = 0;;) {
        if (q == 0.0) q = 1.0e-300;
        c += q < 0.0;
        if (++i >= n) break;
        q = d[i] - x - e[i - 1] * e[i - 1] / q;
    }
    return c;
}

// k-th smallest eigenvalue (k = 0, 1, ...) of the n x n symmetric
// tridiagonal (d, e) by multisection: the 64 lanes of the calling wavefront
// evaluate Sturm counts at 64 interior points of the bracket, which shrinks
// 65-fold per round, from the Gershgorin interval down to the last bits.
// Every lane of the wavefront must call it; all return the same value.
__device__ __attribute__((always_inline)) double tridiag_kth(int n, const double *d, const double *e, int k)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
    }
    const int lane = threadIdx.x & 63;
    for (int round = 0; round < 16; ++round) {
        const double w = hi - lo;
        if (!(w > 4.0e-16 * fmax(fabs(lo), fabs(hi)) + 1.0e-300)) break;
        const double x = lo + w * (double)(lane + 1) / 65.0;
        const unsigned long long above = __ballot(sturm_count(n, d, e, x) > k);
        const int first = above ? __ffsll((long long)above) - 1 : 64;  // bracket (x_first-1, x_first]
        const double nlo = first == 0 ? lo : lo + w * (double)first / 65.0;
        const double nhi = first == 64 ? hi : lo + w * (double)(first + 1) / 65.0;
        lo = nlo;
        hi = nhi;
    }
    return 0.5 * (lo + hi);
}

// Full re-orthogonalisation of the Lanczos residual T against q_0 .. q_k
// with the vectors in LDS: the free atoms' components of q_0 .. q_k and T
// are loaded once (all loads in flight together), then classical
// Gram-Schmidt runs twice, each wavefront taking every fourth projection
// (a wavefront sum, no block barrier per projection), and T goes back to
// global memory.  Element e of a vector is component e / nact of free atom
// e % nact.  Modified Gram-Schmidt in global memory, with k + 1 block
// reductions in a row, stays for vectors too large for the LDS budget.
constexpr int kLzLds = 8192;  // doubles of LDS for q_0 .. q_k and T
__device__ __attribute__((always_inline)) void reorthogonalise_lds(const Ctx &C_, int k)
{
    __shared__ double lq[kLzLds];
    __shared__ double dj[AKMC_LZ_MAX];
    const int na = C_.nact, ne = 3 * na;
    const size_t A3 = 3 * (size_t)C_.A;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // a thread's elements e: their free atom and component once, then the
    // k + 2 vectors' values, kLzIlp loads in flight before their LDS stores
    constexpr int kLzIlp = 4;
    for (int e = threadIdx.x; e < ne; e += kBlock) {
        const int c = e / na, t = e - c * na;
        const size_t o = (size_t)c * C_.A + C_.s.act[C_.off + t];
        for (int j0 = 0; j0 < k + 2; j0 += kLzIlp) {
            double v[kLzIlp];
#pragma unroll
            for (int u = 0; u < kLzIlp; ++u) {
                const int j = j0 + u;
                if (j < k + 2) v[u] = j <= k ? C_.s.Q[j * A3 + o] : C_.s.T[o];
            }
#pragma unroll
            for (int u = 0; u < kLzIlp; ++u)
                if (j0 + u < k + 2) lq[(j0 + u) * ne + e] = v[u];
        }
    }
    double *T = lq + (k + 1) * ne;
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
        for (int j = wave; j <= k; j += kWaves) {
            const double *q = lq + j * ne;
            double d = 0.0;
            for (int e = lane; e < ne; e += 64) d += T[e] * q[e];
            for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o);
            if (lane == 0) dj[j] = d;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < ne; e += kBlock) {
            double r = T[e];
            for (int j = 0; j <= k; ++j) r -= dj[j] * lq[j * ne + e];
            T[e] = r;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < ne; e += kBlock) {
        const int c = e / na, t = e - c * na;
        V3(C_.s.T, c, C_.s.act[C_.off + t]) = T[e];
    }
    __syncthreads();
}

// FP64 check of a converged saddle: max force on the free atoms, energy,
// and the two lowest Ritz values of the finite-difference Hessian from a
// Lanczos run started on the dimer mode (full re-orthogonalisation).  The
// Hessian products are central differences (two force calls per vector):
// the one-sided product of the reference's ART Lanczos carries an
// O(delta E''') bias that, for vectors concentrated on stiff contacts, can
// put a Ritz value below the true lowest curvature.
__device__ __attribute__((always_inline)) void verify_logic(const Ctx &C_, WalkerState &w, double *sh)
{
    const double delta = C_.s.verify_delta;
    const int nl = C_.s.verify_nl < AKMC_LZ_MAX ? C_.s.verify_nl : AKMC_LZ_MAX;
    const size_t A3 = 3 * (size_t)C_.A;
    double *lz = C_.s.lz + (size_t)blockIdx.x * 3 * AKMC_LZ_MAX;
    w.fcalls += 1;
    if (w.state == ST_V_F0) {
        double acc[1] = {0.0};
        double fm2 = 0.0;
        FOR_ACTIVE(i)
        {
            double q = 0.0;
            for (int c = 0; c < 3; ++c) {
                double F = C_.f(c, i);
                V3(C_.s.F0, c, i) = F;
                q += F * F;
                const double n = V3(C_.s.N, c, i);
                acc[0] += n * n;
            }
            fm2 = fmax(fm2, q);
        }
        fm2 = bmax(fm2, sh);
        bsum<1>(acc, sh);
        const double inorm = 1.0 / sqrt(acc[0]);
        FOR_ACTIVE(i)
        {
            for (int c = 0; c < 3; ++c) V3(C_.s.Q, c, i) = V3(C_.s.N, c, i) * inorm;
        }
        w.fmax64 = sqrt(fm2);
        w.e64 = C_.b.out[blockIdx.x].total;
        w.lk = 0;
        w.state = ST_V_LZ;
        __syncthreads();
        write_probe(C_, C_.s.Q, delta);
        return;
    }
    if (w.state == ST_V_LZ) {
        // F(X + delta q_k) kept in F1 while the force at X - delta q_k is taken
        FOR_ACTIVE(i)
        {
            for (int c = 0; c < 3; ++c) V3(C_.s.F1, c, i) = C_.f(c, i);
        }
        w.state = ST_V_LZM;
        __syncthreads();
        write_probe(C_, C_.s.Q + w.lk * A3, -delta);
        return;
    }
    const int k = w.lk;
    const double *Qk = C_.s.Q + k * A3;
    double a1[1] = {0.0};
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            // central difference: H q = -(F(X + d q) - F(X - d q)) / 2d + O(d^2)
            double u = -(V3(C_.s.F1, c, i) - C_.f(c, i)) / (2.0 * delta);
            V3(C_.s.T, c, i) = u;
            a1[0] += V3(Qk, c, i) * u;
        }
    }
    bsum<1>(a1, sh);
    const double alpha = a1[0];
    const double bprev = k > 0 ? lz[AKMC_LZ_MAX + k - 1] : 0.0;
    const double *Qp = k > 0 ? C_.s.Q + (k - 1) * A3 : nullptr;
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            double r = V3(C_.s.T, c, i) - alpha * V3(Qk, c, i);
            if (Qp) r -= bprev * V3(Qp, c, i);
            V3(C_.s.T, c, i) = r;
        }
    }
    // full re-orthogonalisation against q_0 .. q_k
    if (3 * C_.nact * (k + 2) <= kLzLds) {
        __syncthreads();  // the update above is in T
        reorthogonalise_lds(C_, k);
    } else {
        for (int j = 0; j <= k; ++j) {
            const double *Qj = C_.s.Q + j * A3;
            double dj[1] = {0.0};
            FOR_ACTIVE(i)
            {
                for (int c = 0; c < 3; ++c) dj[0] += V3(C_.s.T, c, i) * V3(Qj, c, i);
            }
            bsum<1>(dj, sh);
            FOR_ACTIVE(i)
            {
                for (int c = 0; c < 3; ++c) V3(C_.s.T, c, i) -= dj[0] * V3(Qj, c, i);
            }
        }
    }
    double bb[1] = {0.0};
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) bb[0] += V3(C_.s.T, c, i) * V3(C_.s.T, c, i);
    }
    bsum<1>(bb, sh);
    const double beta = sqrt(bb[0]);
    if (threadIdx.x == 0) {
        lz[k] = alpha;
        lz[AKMC_LZ_MAX + k] = beta;
    }
    w.lk = k + 1;
    if (w.lk >= nl || beta < 1.0e-12) {
        // the two lowest Ritz values, one wavefront each, the tridiagonal in
        // LDS (every lane reads the same entries)
        __shared__ double ev[2], td[AKMC_LZ_MAX], te[AKMC_LZ_MAX];
        const int n = w.lk;
        if (threadIdx.x < (unsigned)n) {
            const int t = threadIdx.x;
            td[t] = t < k ? lz[t] : alpha;
            te[t] = t < k ? lz[AKMC_LZ_MAX + t] : 0.0;
        }
        __syncthreads();
        if (threadIdx.x < 128) {
            const int which = threadIdx.x >> 6;
            const double l = which == 0 || n > 1 ? tridiag_kth(n, td, te, which) : 1.0e300;
            if ((threadIdx.x & 63) == 0) ev[which] = l;
        }
        __syncthreads();
        w.lam1 = ev[0];
        w.lam2 = ev[1];
        w.verified_state = 1;
        end_converged(C_, w, sh);
        return;
    }
    double *Qn = C_.s.Q + (size_t)(k + 1) * A3;
    const double ib = 1.0 / beta;
    FOR_ACTIVE(i)
    {
        double pr[3];
        for (int c = 0; c < 3; ++c) {
            double q = V3(C_.s.T, c, i) * ib;
            V3(Qn, c, i) = q;
            pr[c] = V3(C_.s.X, c, i) + delta * q;
        }
        C_.probe(i, pr[0], pr[1], pr[2]);
    }
    w.state = ST_V_LZ;
}

// ------------------------------------------------------------------- ART

// lowest eigenvector of the n x n symmetric tridiagonal (d, e) for the
// eigenvalue lam by inverse iteration (three sweeps); single thread
__device__ __attribute__((always_inline)) void tridiag_lowest_vector(int n, const double *d, const double *e, double lam, double *y)
{
    double scale = 0.0;
    for (int i = 0; i < n; ++i) scale = fmax(scale, fabs(d[i]) + (i < n - 1 ? fabs(e[i]) : 0.0));
    const double sigma = lam - 1.0e-10 * (scale > 0.0 ? scale : 1.0);
    double c[AKMC_LZ_MAX], z[AKMC_LZ_MAX];
    for (int i = 0; i < n; ++i) y[i] = 1.0;
    for (int sweep = 0; sweep < 3; ++sweep) {
        // Thomas algorithm on (T - sigma I) z = y
        double piv = d[0] - sigma;
        if (fabs(piv) < 1.0e-300) piv = 1.0e-300;
        c[0] = n > 1 ? e[0] / piv : 0.0;
        z[0] = y[0] / piv;
        for (int i = 1; i < n; ++i) {
            piv = d[i] - sigma - e[i - 1] * c[i - 1];
            if (fabs(piv) < 1.0e-300) piv = 1.0e-300;
            c[i] = i < n - 1 ? e[i] / piv : 0.0;
            z[i] = (y[i] - e[i - 1] * z[i - 1]) / piv;
        }
        for (int i = n - 2; i >= 0; --i) z[i] -= c[i] * z[i + 1];
        double nn = 0.0;
        for (int i = 0; i < n; ++i) nn += z[i] * z[i];
        nn = 1.0 / sqrt(nn);
        for (int i = 0; i < n; ++i) y[i] = z[i] * nn;
    }
}

__device__ __attribute__((always_inline)) void art_fire_pre(const Ctx &C_, WalkerState &w, double *sh);
__device__ __attribute__((always_inline)) void enter_verify_or_finish(const Ctx &C_, WalkerState &w, double curvature, double *sh);

// start of ART::calculateLanczosForce (optim/ART.cpp:159-190): force at x,
// q_0 = r / |r| with r the current lowest-mode estimate
__device__ __attribute__((always_inline)) void art_lanczos_begin(const Ctx &C_, WalkerState &w, double *sh)
{
    const ARTParamsDev &p = C_.s.ap;
    double nn[1] = {0.0};
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            V3(C_.s.F0, c, i) = C_.f(c, i);
            double r = V3(C_.s.N, c, i);
            nn[0] += r * r;
        }
    }
    bsum<1>(nn, sh);
    w.E0 = C_.b.out[blockIdx.x].total;
    const double beta = sqrt(nn[0]);
    w.lk = 0;
    w.lbeta = beta;
    w.lconv = 0;
    w.lam = 1.0e6;  // lowestEigenValue before the first diagonalisation (ART.cpp:157)
    FOR_ACTIVE(i)
    {
        double pr[3];
        for (int c = 0; c < 3; ++c) {
            double q = V3(C_.s.N, c, i) / beta;
            V3(C_.s.Q, c, i) = q;
            pr[c] = V3(C_.s.X, c, i) + p.deltaX * q;
        }
        C_.probe(i, pr[0], pr[1], pr[2]);
    }
    w.state = ST_A_LZ;
}

// ART after a Lanczos evaluation: velocity half kick and convergence test
// of the previous iteration (optim/ART.cpp:660-720), or entry into the loop
__device__ __attribute__((always_inline)) void art_after_lanczos(const Ctx &C_, WalkerState &w, double *sh)
{
    const ARTParamsDev &p = C_.s.ap;
    if (!w.inloop) {
        w.inloop = 1;
        w.half = 0.5 * w.dt * p.timeStepUnit;
        art_fire_pre(C_, w, sh);
        return;
    }
    const bool perp = w.lam >= p.maxEigenvalue;
    const double dot = w.dotfv;
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            double v = V3(C_.s.N, c, i);
            double F = perp ? -dot * v : V3(C_.s.F0, c, i) - 2.0 * dot * v;
            V3(C_.s.V, c, i) += w.half * F;
        }
    }
    if (w.it > 5 && ((w.fnpar + w.fnperp) < p.normTol || (w.fmpar + w.fmperp) < p.forceTol)) {
        w.status = (w.lam <= p.maxEigenvalue) ? 0 : 1;
        if (w.status == 0) enter_verify_or_finish(C_, w, w.lam, sh);
        else finish(C_, w);
        return;
    }
    w.half = 0.5 * w.dt * p.timeStepUnit;
    if (fabs(w.half) < 1.0e-10) {
        w.status = 1;
        finish(C_, w);
        return;
    }
    art_fire_pre(C_, w, sh);
}

// one Lanczos step (optim/ART.cpp:191-283); at the end the Ritz vector,
// F_par = F - 2(F.v)v and F_perp = -(F.v)v (ART.cpp:318-349)
__device__ __attribute__((always_inline)) void art_lanczos_step(const Ctx &C_, WalkerState &w, double *sh)
{
    const ARTParamsDev &p = C_.s.ap;
    const size_t A3 = 3 * (size_t)C_.A;
    double *lz = C_.s.lz + (size_t)blockIdx.x * 3 * AKMC_LZ_MAX;
    const int i = w.lk;
    const int nl = p.numLanczosVectors;
    const double *Qi = C_.s.Q + i * A3;
    const double *Qp = i > 0 ? C_.s.Q + (i - 1) * A3 : nullptr;
    const double beta = w.lbeta;
    double a1[1] = {0.0};
    FOR_ACTIVE(k)
    {
        for (int c = 0; c < 3; ++c) {
            double u = -(C_.f(c, k) - V3(C_.s.F0, c, k)) / p.deltaX;
            double r = Qp ? u - beta * V3(Qp, c, k) : u;
            V3(C_.s.T, c, k) = r;
            a1[0] += V3(Qi, c, k) * r;
        }
    }
    bsum<1>(a1, sh);
    const double alpha = a1[0];
    double bb[1] = {0.0};
    FOR_ACTIVE(k)
    {
        for (int c = 0; c < 3; ++c) {
            double r = V3(C_.s.T, c, k) - alpha * V3(Qi, c, k);
            V3(C_.s.T, c, k) = r;
            bb[0] += r * r;
        }
    }
    bsum<1>(bb, sh);
    const double bnew = sqrt(bb[0]);
    __shared__ double evs[1];
    if (threadIdx.x == 0) {
        lz[i] = alpha;
        if (i > 0) lz[AKMC_LZ_MAX + i - 1] = beta;
    }
    __shared__ double td[AKMC_LZ_MAX], te[AKMC_LZ_MAX];  // the tridiagonal, in LDS
    if (i > 0 && threadIdx.x <= (unsigned)i) {
        const int t = threadIdx.x;
        td[t] = t < i ? lz[t] : alpha;
        te[t] = t < i - 1 ? lz[AKMC_LZ_MAX + t] : (t == i - 1 ? beta : 0.0);
    }
    __syncthreads();
    if (i == 0) {
        if (threadIdx.x == 0) evs[0] = alpha;
    } else if (threadIdx.x < 64) {  // one wavefront: the lowest Ritz value
        const double l1 = tridiag_kth(i + 1, td, te, 0);
        if (threadIdx.x == 0) evs[0] = l1;
    }
    __syncthreads();
    const double evnew = evs[0];
    if (i == 0) {
        w.lamold = evnew;
    } else {
        double delta = fabs((evnew - w.lamold) / w.lamold);
        w.lamold = evnew;
        w.lam = evnew;
        if (delta < p.eigValTol) w.lconv = 1;
    }
    const int next = i + 1;
    if (next < nl && !w.lconv) {
        double *Qn = C_.s.Q + next * A3;
        const double ib = 1.0 / bnew;
        FOR_ACTIVE(k)
        {
            double pr[3];
            for (int c = 0; c < 3; ++c) {
                double q = V3(C_.s.T, c, k) * ib;
                V3(Qn, c, k) = q;
                pr[c] = V3(C_.s.X, c, k) + p.deltaX * q;
            }
            C_.probe(k, pr[0], pr[1], pr[2]);
        }
        w.lk = next;
        w.lbeta = bnew;
        return;
    }
    // Ritz vector of the last diagonalised T (size next)
    if (threadIdx.x == 0) {
        double d[AKMC_LZ_MAX], e[AKMC_LZ_MAX];
        for (int t = 0; t < next; ++t) {
            d[t] = lz[t];
            e[t] = t < next - 1 ? lz[AKMC_LZ_MAX + t] : 0.0;
        }
        tridiag_lowest_vector(next, d, e, evnew, lz + 2 * AKMC_LZ_MAX);
    }
    __syncthreads();
    const double *y = lz + 2 * AKMC_LZ_MAX;
    double nn[1] = {0.0};
    FOR_ACTIVE(k)
    {
        for (int c = 0; c < 3; ++c) {
            double v = 0.0;
            for (int t = 0; t < next; ++t) v += y[t] * V3(C_.s.Q + t * A3, c, k);
            V3(C_.s.N, c, k) = v;
            nn[0] += v * v;
        }
    }
    bsum<1>(nn, sh);
    const double inv = 1.0 / sqrt(nn[0]);
    double dd[1] = {0.0};
    FOR_ACTIVE(k)
    {
        for (int c = 0; c < 3; ++c) {
            double v = V3(C_.s.N, c, k) * inv;
            V3(C_.s.N, c, k) = v;
            dd[0] += V3(C_.s.F0, c, k) * v;
        }
    }
    bsum<1>(dd, sh);
    w.dotfv = dd[0];
    w.C = w.lam;  // reported curvature: the Lanczos lowest eigenvalue
    art_after_lanczos(C_, w, sh);
}

// the FIRE part of one ART iteration before the next Lanczos evaluation
// (optim/ART.cpp:549-656)
__device__ __attribute__((always_inline)) void art_fire_pre(const Ctx &C_, WalkerState &w, double *sh)
{
    const ARTParamsDev &p = C_.s.ap;
    if (w.it > p.maxParIter) {
        w.status = 1;
        finish(C_, w);
        return;
    }
    w.it += 1;
    const bool perp = w.lam >= p.maxEigenvalue;
    const double dot = w.dotfv;
    double acc[3] = {0, 0, 0};  // |F_par|^2, |F_perp|^2, |v|^2
    double mx[3] = {0, 0, 0};
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            double nv = V3(C_.s.N, c, i);
            double fpar = V3(C_.s.F0, c, i) - 2.0 * dot * nv, fperp = -dot * nv;
            double v = V3(C_.s.V, c, i) + w.half * (perp ? fperp : fpar * p.parDamp);
            V3(C_.s.V, c, i) = v;
            acc[0] += fpar * fpar;
            acc[1] += fperp * fperp;
            acc[2] += v * v;
            mx[0] = fmax(mx[0], fpar * fpar);
            mx[1] = fmax(mx[1], fperp * fperp);
        }
    }
    bsum<3>(acc, sh);
    mx[0] = bmax(mx[0], sh);
    mx[1] = bmax(mx[1], sh);
    w.fnpar = sqrt(acc[0]);
    w.fnperp = sqrt(acc[1]);
    w.fmpar = sqrt(mx[0]);
    w.fmperp = sqrt(mx[1]);
    const double vn = sqrt(acc[2]);
    const double fsel = perp ? w.fnperp : w.fnpar;
    double P[1] = {0.0};
    FOR_ACTIVE(i)
    {
        for (int c = 0; c < 3; ++c) {
            double nv = V3(C_.s.N, c, i);
            double F = perp ? -dot * nv : V3(C_.s.F0, c, i) - 2.0 * dot * nv;
            P[0] += F * V3(C_.s.V, c, i);
        }
    }
    bsum<1>(P, sh);
    const double al = w.alpha;
    const bool down = P[0] > 0;
    if (down) {
        if (w.nstep > p.N_min) {
            w.dt = fmin(w.dt * p.stepIncrease, p.maxTimeStep);
            w.alpha *= p.alphaDecrease;
        }
        w.nstep += 1;
    } else {
        w.nstep = 0;
        w.dt *= p.stepDecrease;
        w.alpha = p.alphaStart;
    }
    const double dt = w.dt;
    FOR_ACTIVE(i)
    {
        const bool fr = C_.s.frozen[i] != 0;
        for (int c = 0; c < 3; ++c) {
            double nv = V3(C_.s.N, c, i);
            double F = perp ? -dot * nv : V3(C_.s.F0, c, i) - 2.0 * dot * nv;
            double v = down ? (1.0 - al) * V3(C_.s.V, c, i) + al * F * vn / fsel : 0.0;
            V3(C_.s.V, c, i) = v;
            if (!fr) V3(C_.s.X, c, i) += dt * v;
        }
    }
    w.state = ST_A_L0;
    __syncthreads();
    write_probe(C_, nullptr, 0.0);
}

__device__ __attribute__((always_inline)) void art_logic(const Ctx &C_, WalkerState &w, double *sh)
{
    w.fcalls += 1;
    if (w.state == ST_A_L0) art_lanczos_begin(C_, w, sh);
    else art_lanczos_step(C_, w, sh);
}

__global__ void __launch_bounds__(kBlock) k_search_logic(BatchDev b, SearchDev s)
{
    __shared__ double sh[4 * kBlock];
    __shared__ WalkerState ws;
    const int wid = blockIdx.x;
    if (threadIdx.x == 0) {
        ws = s.ws[wid];
        // this step's FP32 pair kernels evaluated the walker iff it is
        // still marked FP32 here (the logic below may change the mark)
        if (b.active[wid] == 1) s.evals32[wid] += 1;
    }
    __syncthreads();
    if (ws.kind == WK_IDLE || ws.state == ST_DONE) return;
    const WalkerDev W = b.walkers[wid];
    Ctx C_{b, s, W.off, W.n, b.A, s.nact[wid]};
    // every thread keeps a private copy of the walker scalars; all threads
    // take the same branches (the scalars only change through block-wide
    // reductions), thread 0 writes them back
    WalkerState w = ws;
    if (w.kind == WK_RELAX || w.kind == WK_BASIN2) relax_logic(C_, w, sh);
    else if (w.kind == WK_DIMER) dimer_logic(C_, w, sh);
    else if (w.kind == WK_VERIFY) verify_logic(C_, w, sh);
    else if (w.kind == WK_ART) art_logic(C_, w, sh);
    if (threadIdx.x == 0) s.ws[wid] = w;
}

// probes of freshly started walkers (no force call consumed yet): X for
// relax and the dimer centre.  Walkers already under way keep the probe
// their last logic step wrote (a dimer end, a Lanczos displacement).
__global__ void __launch_bounds__(kBlock) k_probe_init(BatchDev b, SearchDev s)
{
    const int wid = blockIdx.x;
    const WalkerState w = s.ws[wid];
    if (w.kind == WK_IDLE || w.state == ST_DONE || w.fcalls != 0) return;
    const WalkerDev W = b.walkers[wid];
    Ctx C_{b, s, W.off, W.n, b.A, s.nact[wid]};
    FOR_ATOMS(i) C_.probe(i, V3(s.X, 0, i), V3(s.X, 1, i), V3(s.X, 2, i));
}

#if AKMC_LOGIC_BLOCK == 256
// scatter of packed walker starts (one workgroup per walker): mode N,
// zero velocity, positions X when given, the frozen mask, the free-atom
// list and its length
__global__ void __launch_bounds__(kBlock) k_load_walkers(BatchDev b, SearchDev s, const LoadItem *items,
                                                          const double *D, const int *I)
{
    const LoadItem it = items[blockIdx.x];
    const WalkerDev W = b.walkers[it.walker];
    const int n = W.n, off = W.off;
    const size_t A = b.A;
    const double *d = D + it.dbl;
    const int *iw = I + it.ints;
    int *frozen = const_cast<int *>(s.frozen);
    int *act = const_cast<int *>(s.act);
    for (int i = threadIdx.x; i < n; i += kBlock) {
        for (int c = 0; c < 3; ++c) {
            s.N[c * A + off + i] = d[(size_t)c * n + i];
            s.V[c * A + off + i] = 0.0;
            if (it.has_x) s.X[c * A + off + i] = d[(size_t)(3 + c) * n + i];
        }
        frozen[off + i] = iw[i];
    }
    for (int k = threadIdx.x; k < it.nact; k += kBlock) act[off + k] = iw[n + k];
    if (threadIdx.x == 0) const_cast<int *>(s.nact)[it.walker] = it.nact;
}

#endif
}  // namespace

#if AKMC_LOGIC_BLOCK == 256
void akmc_launch_load_walkers(const BatchDev &b, const SearchDev &s, const LoadItem *items, int nitems,
                              const double *D, const int *I, hipStream_t st)
{
    if (nitems > 0) k_load_walkers<<<nitems, kBlock, 0, st>>>(b, s, items, D, I);
}

// the listed walkers' x, y, z slices packed walker after walker (out +
// dst[k]: x[n], y[n], z[n] of walker ws[k]); one workgroup per walker
__global__ void __launch_bounds__(kBlock) k_gather_walkers(const WalkerDev *walkers, const double *sx,
                                                           const double *sy, const double *sz, const int *ws,
                                                           const long long *dst, double *out)
{
    const WalkerDev W = walkers[ws[blockIdx.x]];
    double *o = out + dst[blockIdx.x];
    for (int a = threadIdx.x; a < W.n; a += kBlock) {
        o[a] = sx[W.off + a];
        o[W.n + a] = sy[W.off + a];
        o[2 * W.n + a] = sz[W.off + a];
    }
}

void akmc_launch_gather_walkers(const WalkerDev *walkers, const double *const src[3], const int *ws,
                                const long long *dst, int nw, double *out, hipStream_t st)
{
    if (nw > 0) k_gather_walkers<<<nw, kBlock, 0, st>>>(walkers, src[0], src[1], src[2], ws, dst, out);
}

void akmc_launch_probe_init(const BatchDev &b, const SearchDev &s, hipStream_t st)
{
    k_probe_init<<<b.W, kBlock, 0, st>>>(b, s);
}

__global__ void k_start_walkers(BatchDev b, SearchDev s, const StartItem *items, int n, int *zflag, int *harvested)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const StartItem &it = items[k];
    s.ws[it.walker] = it.ws;
    if (zflag) zflag[it.walker] = it.zflag;
    if (harvested) harvested[it.walker] = 0;
    b.rebuild[it.walker] = 1;
    b.active[it.walker] = it.active;
}

void akmc_launch_start_walkers(const BatchDev &b, const SearchDev &s, const StartItem *items, int nitems, int *zflag,
                               int *harvested, hipStream_t st)
{
    if (nitems > 0)
        k_start_walkers<<<(nitems + kBlock - 1) / kBlock, kBlock, 0, st>>>(b, s, items, nitems, zflag, harvested);
}

// one workgroup per walker
__global__ void __launch_bounds__(kBlock) k_harvest_pack(BatchDev b, SearchDev s, int *harvested, double *out)
{
    const int w = blockIdx.x;
    const WalkerState &ws = s.ws[w];
    if (ws.kind == WK_IDLE || ws.state != ST_DONE || harvested[w]) return;
    const WalkerDev W = b.walkers[w];
    const size_t A = b.A;
    for (int i = W.off + (int)threadIdx.x; i < W.off + W.n; i += kBlock)
        for (int c = 0; c < 3; ++c) out[(size_t)c * A + i] = s.X[(size_t)c * A + i];
    if (threadIdx.x == 0) harvested[w] = 1;
}

void akmc_launch_harvest_pack(const BatchDev &b, const SearchDev &s, int *harvested, double *out, hipStream_t st)
{
    k_harvest_pack<<<b.W, kBlock, 0, st>>>(b, s, harvested, out);
}

void akmc_launch_search_logic(const BatchDev &b, const SearchDev &s, hipStream_t st)
{
    k_search_logic<<<b.W, kBlock, 0, st>>>(b, s);
}
#else
void akmc_launch_search_logic_wide(const BatchDev &b, const SearchDev &s, hipStream_t st)
{
    k_search_logic<<<b.W, kBlock, 0, st>>>(b, s);
}
#endif
