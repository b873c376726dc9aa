// Walker state machines for gfx950: FIRE relax (optim/Relax.cpp:44-315)
// and the dimer saddle search (rotation + FIRE translation), advanced by
// one logic kernel after every batched force evaluation.
//
// One workgroup owns one walker.  Every vector operation of the walker is a
// strided loop of its 256 threads over the walker's atoms followed by a
// fixed-order LDS tree reduction, so a walker's trajectory is bitwise the
// same whether it runs alone or in a batch of hundreds.  Vectors are FP64
// [3][A] arrays (component c of atom i at c*A + i).
#include "akmc_internal.h"

#include <math.h>

namespace {

constexpr int kBlock = 256;

template <int NV>
__device__ inline void bsum(double (&v)[NV], double *sh)
{
    for (int c = 0; c < NV; ++c) sh[c * kBlock + threadIdx.x] = v[c];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o)
            for (int c = 0; c < NV; ++c) sh[c * kBlock + threadIdx.x] += sh[c * kBlock + threadIdx.x + o];
        __syncthreads();
    }
    for (int c = 0; c < NV; ++c) v[c] = sh[c * kBlock];
    __syncthreads();
}

__device__ inline double bmax(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    v = sh[0];
    __syncthreads();
    return v;
}

struct Ctx {
    const BatchDev &b;
    const SearchDev &s;
    int off, n, A;
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

#define FOR_ATOMS(i) for (int i = C_.off + (int)threadIdx.x; i < C_.off + C_.n; i += kBlock)
#define V3(arr, c, i) (arr)[(size_t)(c) * C_.A + (i)]

// probe = X + a * D (D a [3][A] vector) or X when D is null
__device__ inline void write_probe(const Ctx &C_, const double *D, double a)
{
    FOR_ATOMS(i)
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

// ------------------------------------------------------------------ relax

// the part of a FIRE iteration before its force call (Relax.cpp:125-193,
// semi-implicit Euler)
__device__ void relax_pre(const Ctx &C_, WalkerState &w, double *sh)
{
    const RelaxParamsDev &p = C_.s.rp;
    double acc[3] = {0, 0, 0};  // P, |F|^2, |V|^2
    FOR_ATOMS(i)
    {
        for (int c = 0; c < 3; ++c) {
            double F = V3(C_.s.F0, c, i), v = V3(C_.s.V, c, i);
            acc[0] += F * v;
            acc[1] += F * F;
            acc[2] += v * v;
        }
    }
    bsum<3>(acc, sh);
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
    } else {
        w.nstep = 0;
        w.dt *= p.stepDecrease;
        w.alpha = p.alphaStart;
    }
    double dt = w.dt, fact = dt * p.timeStepUnit;
    FOR_ATOMS(i)
    {
        bool fr = C_.s.frozen[i] != 0;
        for (int c = 0; c < 3; ++c) {
            double v = V3(C_.s.V, c, i), F = V3(C_.s.F0, c, i);
            if (down) {
                if (!fr) v = oma * v + af * F;
            } else {
                v = 0.0;
            }
            if (!fr) {
                V3(C_.s.X, c, i) += dt * v;
                v += fact * F;
            }
            V3(C_.s.V, c, i) = v;
        }
    }
    __syncthreads();
    write_probe(C_, nullptr, 0.0);
}

__device__ void relax_logic(const Ctx &C_, WalkerState &w, double *sh)
{
    const RelaxParamsDev &p = C_.s.rp;
    // forces at X (frozen atoms zeroed, Relax.cpp:230)
    FOR_ATOMS(i)
    {
        for (int c = 0; c < 3; ++c) V3(C_.s.F0, c, i) = C_.f(c, i);
    }
    w.fcalls += 1;
    w.E_last = C_.b.out[blockIdx.x].total;
    w.E0 = w.E_last;
    __syncthreads();
    if (w.state == ST_R_INIT) {
        w.it = 1;
        w.state = ST_R_STEP;
        relax_pre(C_, w, sh);
        return;
    }
    // centre-of-mass velocity removal over all atoms (Relax.cpp:964-996)
    double m[3] = {0, 0, 0};
    FOR_ATOMS(i)
    {
        for (int c = 0; c < 3; ++c) m[c] += V3(C_.s.V, c, i);
    }
    bsum<3>(m, sh);
    double fn2 = 0.0, fm2 = 0.0;
    FOR_ATOMS(i)
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
        w.status = 0;
        w.state = ST_DONE;
        write_probe(C_, nullptr, 0.0);
        return;
    }
    if (w.it >= p.maxIter) {
        w.status = 1;
        w.state = ST_DONE;
        write_probe(C_, nullptr, 0.0);
        return;
    }
    w.it += 1;
    relax_pre(C_, w, sh);
}

// ------------------------------------------------------------------ dimer

__device__ void dimer_translate(const Ctx &C_, WalkerState &w, double *sh);

// rotation bookkeeping: decide between another trial rotation and the
// translation (the rotation loop of the CPU oracle's dimer)
__device__ void dimer_rot_eval(const Ctx &C_, WalkerState &w, double *sh)
{
    const DimerParamsDev &p = C_.s.dp;
    const double D = p.dimerDisplacement;
    double acc[1] = {0.0};
    FOR_ATOMS(i)
    {
        for (int c = 0; c < 3; ++c) acc[0] += (V3(C_.s.F0, c, i) - V3(C_.s.F1, c, i)) / D * V3(C_.s.N, c, i);
    }
    bsum<1>(acc, sh);
    double Cv = acc[0];
    double g2[1] = {0.0};
    FOR_ATOMS(i)
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
    FOR_ATOMS(i)
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
__device__ void dimer_translate(const Ctx &C_, WalkerState &w, double *sh)
{
    const DimerParamsDev &p = C_.s.dp;
    const double D = p.dimerDisplacement;
    double acc[2] = {0.0, 0.0};  // C*D, F0.n
    double fm2 = 0.0;
    FOR_ATOMS(i)
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
        w.state = ST_DONE;
        write_probe(C_, nullptr, 0.0);
        return;
    }
    const bool convex = Cv > 0.0;
    double a3[3] = {0, 0, 0};  // Fe.V, |V|^2, |Fe|^2
    FOR_ATOMS(i)
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
    FOR_ATOMS(i)
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
    FOR_ATOMS(i)
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
        w.state = ST_DONE;
        __syncthreads();
        write_probe(C_, nullptr, 0.0);
        return;
    }
    w.state = ST_D_F0;
    __syncthreads();
    write_probe(C_, nullptr, 0.0);
}

__device__ void dimer_logic(const Ctx &C_, WalkerState &w, double *sh)
{
    const DimerParamsDev &p = C_.s.dp;
    const double D = p.dimerDisplacement;
    w.fcalls += 1;
    if (w.state == ST_D_F0) {
        FOR_ATOMS(i)
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
        FOR_ATOMS(i)
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
    FOR_ATOMS(i)
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
    FOR_ATOMS(i)
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
    FOR_ATOMS(i)
    {
        for (int c = 0; c < 3; ++c) V3(C_.s.N, c, i) *= inv;
    }
    w.nrot += 1;
    __syncthreads();
    dimer_rot_eval(C_, w, sh);
}

__global__ void __launch_bounds__(kBlock) k_search_logic(BatchDev b, SearchDev s)
{
    __shared__ double sh[4 * kBlock];
    __shared__ WalkerState ws;
    const int wid = blockIdx.x;
    if (threadIdx.x == 0) ws = s.ws[wid];
    __syncthreads();
    if (ws.kind == WK_IDLE || ws.state == ST_DONE) return;
    const WalkerDev W = b.walkers[wid];
    Ctx C_{b, s, W.off, W.n, b.A};
    // every thread keeps a private copy of the walker scalars; all threads
    // take the same branches (the scalars only change through block-wide
    // reductions), thread 0 writes them back
    WalkerState w = ws;
    if (w.kind == WK_RELAX) relax_logic(C_, w, sh);
    else if (w.kind == WK_DIMER) dimer_logic(C_, w, sh);
    if (threadIdx.x == 0) s.ws[wid] = w;
}

// probes of freshly started walkers: X for relax and the dimer centre
__global__ void __launch_bounds__(kBlock) k_probe_init(BatchDev b, SearchDev s)
{
    const int wid = blockIdx.x;
    const WalkerState w = s.ws[wid];
    if (w.kind == WK_IDLE || w.state == ST_DONE) return;
    const WalkerDev W = b.walkers[wid];
    Ctx C_{b, s, W.off, W.n, b.A};
    write_probe(C_, nullptr, 0.0);
}

}  // namespace

void akmc_launch_search_logic(const BatchDev &b, const SearchDev &s, hipStream_t st)
{
    k_search_logic<<<b.W, kBlock, 0, st>>>(b, s);
}

void akmc_launch_probe_init(const BatchDev &b, const SearchDev &s, hipStream_t st)
{
    k_probe_init<<<b.W, kBlock, 0, st>>>(b, s);
}
