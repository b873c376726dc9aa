// Learned saddle proposal on gfx950: symmetry-function descriptors and the
// MLP 48 -> 128 -> 128 -> 3 (SiLU) with hand-written BF16 MFMA forward and
// backward GEMMs, FP32 accumulation, FP32 master weights and Adam.
//
// GEMM tiles: one wavefront computes a 16 x 16 output tile with
// v_mfma_f32_16x16x16_bf16 (lane l holds A[l%16][4(l/16)..+3],
// B[4(l/16)..+3][l%16] and D[4(l/16)..+3][l%16]); a 256-thread workgroup
// covers 64 rows x 16 columns.  Activations are FP32 in HBM and rounded to
// BF16 (round to nearest even) when loaded into MFMA operands.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "akmc_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kB = 256;

__device__ inline uint16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

enum { EPI_SILU = 0, EPI_LINEAR = 1, EPI_DSILU = 2 };

// C[m][n] = sum_k A[m*lda + k] * Bt[n*ldb + k] (A FP32, Bt BF16 bits);
// K = 16 KS; n < N are stored; A rows 16-byte aligned.  Epilogues:
//   EPI_SILU   Z = acc + bias, stored; Y = silu(Z)
//   EPI_LINEAR Y = acc + bias
//   EPI_DSILU  Y = acc * silu'(aux[m*ldc + n])
// KS = K / 16 k-steps: every operand of the tile is loaded first (one
// 16-byte A load and one 8-byte B load per lane and k-step, all in flight
// together), then the KS MFMAs run back to back in k order.
template <int EPI, int KS>
__global__ void __launch_bounds__(kB) k_gemm(const float *__restrict__ A, int lda, const uint16_t *__restrict__ Bt,
                                             int ldb, int M, int N, const float *__restrict__ bias,
                                             const float *__restrict__ aux, float *__restrict__ Z,
                                             float *__restrict__ Y, int ldc)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = blockIdx.x * 64 + wave * 16, c0 = blockIdx.y * 16;
    const int li = lane & 15, lk = (lane >> 4) * 4;
    const int arow = r0 + li, bcol = c0 + li;
    float4 av[KS];
    bf16x4 bv[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        av[s] = arow < M ? *reinterpret_cast<const float4 *>(A + (size_t)arow * lda + 16 * s + lk)
                         : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (bcol < N) {
            bv[s] = *reinterpret_cast<const bf16x4 *>(Bt + (size_t)bcol * ldb + 16 * s + lk);
        } else {
            bv[s] = bf16x4{0, 0, 0, 0};
        }
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const bf16x4 a = {(short)to_bf16(av[s].x), (short)to_bf16(av[s].y), (short)to_bf16(av[s].z),
                          (short)to_bf16(av[s].w)};
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv[s], acc, 0, 0, 0);
    }
    const int col = c0 + li;
    if (col >= N) return;
    for (int e = 0; e < 4; ++e) {
        const int row = r0 + lk + e;
        if (row >= M) continue;
        float v = acc[e];
        size_t o = (size_t)row * ldc + col;
        if (EPI == EPI_SILU) {
            v += bias[col];
            Z[o] = v;
            Y[o] = v * sigm(v);
        } else if (EPI == EPI_LINEAR) {
            Y[o] = v + bias[col];
        } else {
            float z = aux[o], s = sigm(z);
            Y[o] = v * s * (1.0f + z * (1.0f - s));
        }
    }
}

// split-K weight gradient: P[chunk][k][n] = sum over the chunk's rows m of
// A[m*lda + k] * G[m*ldg + n]  (k < Kd, n < Nd; Kd a multiple of 16)
__global__ void __launch_bounds__(kB) k_wgrad(const float *__restrict__ A, int lda, const float *__restrict__ G,
                                              int ldg, int M, int Kd, int Nd, int rows_per_chunk,
                                              float *__restrict__ P)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i0 = blockIdx.x * 64 + wave * 16, j0 = blockIdx.y * 16, chunk = blockIdx.z;
    const int li = lane & 15, lk = (lane >> 4) * 4;
    const int m_begin = chunk * rows_per_chunk;
    const int m_end = min(M, m_begin + rows_per_chunk);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int ai = i0 + li, bj = j0 + li;
    // four k-steps (64 rows) of operands loaded before their MFMAs
    constexpr int S = 4;
    for (int m00 = m_begin; m00 < m_end; m00 += 16 * S) {
        float av[S][4], gv[S][4];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m00 + 16 * s + lk + e;
                av[s][e] = (m < m_end && ai < Kd) ? A[(size_t)m * lda + ai] : 0.0f;
                gv[s][e] = (m < m_end && bj < Nd) ? G[(size_t)m * ldg + bj] : 0.0f;
            }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (m00 + 16 * s >= m_end) break;
            bf16x4 a, b;
            for (int e = 0; e < 4; ++e) {
                a[e] = (short)to_bf16(av[s][e]);
                b[e] = (short)to_bf16(gv[s][e]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    const int col = j0 + li;
    if (col >= Nd) return;
    for (int e = 0; e < 4; ++e) {
        int row = i0 + lk + e;
        if (row < Kd) P[((size_t)chunk * Kd + row) * Nd + col] = acc[e];
    }
}

// out[i] = sum over chunks of P[chunk][i]  (fixed order)
__global__ void k_sum_chunks(const float *__restrict__ P, int nchunk, int n, float *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int c = 0; c < nchunk; ++c) s += P[(size_t)c * n + i];
    out[i] = s;
}

// column sums of G[m*ldg + n], n < N: one workgroup per column
__global__ void __launch_bounds__(kB) k_colsum(const float *__restrict__ G, int ldg, int M, float *__restrict__ out)
{
    __shared__ float sh[kB];
    const int n = blockIdx.x;
    float s = 0.0f;
    for (int m = threadIdx.x; m < M; m += kB) s += G[(size_t)m * ldg + n];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = sh[0];
}

// dY = 2 (Y - T) / M into a 16-wide padded buffer; per-block loss partials
__global__ void __launch_bounds__(kB) k_mse(const float *__restrict__ Y, const float *__restrict__ T, int M,
                                            float *__restrict__ dY, double *__restrict__ part)
{
    __shared__ double sh[kB];
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (m < M) {
        for (int c = 0; c < 16; ++c) {
            float r = c < 3 ? Y[(size_t)m * 16 + c] - T[(size_t)m * 3 + c] : 0.0f;
            dY[(size_t)m * 16 + c] = 2.0f * r / (float)M;
            l += (double)r * r;
        }
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// BF16 copies of the weights: forward operands (transposed, W^T[n][k]) and
// backward-data operands (row-major W[k][n]), the output layer padded to 16
__global__ void k_weights_bf16(const float *__restrict__ p, uint16_t *__restrict__ w1t, uint16_t *__restrict__ w2t,
                               uint16_t *__restrict__ w3t, uint16_t *__restrict__ w2r, uint16_t *__restrict__ w3r)
{
    const float *W1 = p, *W2 = p + 48 * 128 + 128, *W3 = W2 + 128 * 128 + 128;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 128 * 48) {
        int n = t / 48, k = t % 48;
        w1t[t] = to_bf16(W1[k * 128 + n]);
    }
    if (t < 128 * 128) {
        int n = t / 128, k = t % 128;
        w2t[t] = to_bf16(W2[k * 128 + n]);
        w2r[t] = to_bf16(W2[t]);
    }
    if (t < 16 * 128) {
        int n = t / 128, k = t % 128;
        w3t[t] = n < 3 ? to_bf16(W3[k * 3 + n]) : (uint16_t)0;
        int kr = t / 16, nr = t % 16;
        w3r[t] = nr < 3 ? to_bf16(W3[kr * 3 + nr]) : (uint16_t)0;
    }
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float a = b1 * m1[i] + (1.0f - b1) * gi;
    float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void k_gather_rows(const float *__restrict__ src, int width, const int *__restrict__ idx, int M,
                              float *__restrict__ dst)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * width) return;
    int m = t / width, c = t % width;
    dst[t] = src[(size_t)idx[m] * width + c];
}

// ------------------------------------------------------------ descriptors

__device__ inline float fcut(float r, float rc) { return r < rc ? 0.5f * (__cosf(3.14159265f * r / rc) + 1.0f) : 0.0f; }

// one query per thread: atom i of walker w with vacancy site v and candidate
// mover position u; neighbours from the walker's list at its resident
// positions (minimum image).  48 descriptors: 32 radial G2 (16 over the
// atoms, 8 over the vacancy site and 8 over the mover as pseudo-neighbours
// of their own kinds) and 16 angular G4; the definitions are restated in
// FP64 by the test oracle
__global__ void __launch_bounds__(kB) k_descriptors(BatchDev b, int nq, const int *__restrict__ qatom,
                                                    const double *__restrict__ qvac, const double *__restrict__ qmov,
                                                    float *__restrict__ X)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int i = qatom[q];
    const WalkerDev W = b.walkers[b.wk[i]];
    const double4 pi = b.p4[i];
    const float RC = 7.22f, RCA = 5.0f, ETA_R = 8.0f, ETA_S = 2.0f;
    float g[48];
    for (int t = 0; t < 48; ++t) g[t] = 0.0f;
    const int cnt = b.nnb[i];
    for (int k = 0; k < cnt; ++k) {
        uint32_t e = nbr_entry(b, k, i);
        const double4 pj = b.p4[W.off + int(e & AKMC_NB_JMASK)];
        double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        dx -= W.L[0] * rint(dx * W.invL[0]);
        dy -= W.L[1] * rint(dy * W.invL[1]);
        dz -= W.L[2] * rint(dz * W.invL[2]);
        float r = (float)sqrt(dx * dx + dy * dy + dz * dz);
        if (r < 1.0e-8f || r >= RC) continue;
        float f = fcut(r, RC);
        for (int t = 0; t < 16; ++t) {
            float d = r - (2.0f + 5.0f * t / 15.0f);
            g[t] += __expf(-ETA_R * d * d) * f;
        }
        if (r >= RCA) continue;
        float fa = fcut(r, RCA);
        float ax = (float)dx, ay = (float)dy, az = (float)dz;
        // pairs (j, k) with k later in the list
        for (int k2 = k + 1; k2 < cnt; ++k2) {
            uint32_t e2 = nbr_entry(b, k2, i);
            const double4 pk = b.p4[W.off + int(e2 & AKMC_NB_JMASK)];
            double ex = pk.x - pi.x, ey = pk.y - pi.y, ez = pk.z - pi.z;
            ex -= W.L[0] * rint(ex * W.invL[0]);
            ey -= W.L[1] * rint(ey * W.invL[1]);
            ez -= W.L[2] * rint(ez * W.invL[2]);
            float r2k = (float)sqrt(ex * ex + ey * ey + ez * ez);
            if (r2k < 1.0e-8f || r2k >= RCA) continue;
            float bx = (float)ex - ax, by = (float)ey - ay, bz = (float)ez - az;
            float rjk = sqrtf(bx * bx + by * by + bz * bz);
            if (rjk >= RCA) continue;
            float base = fa * fcut(r2k, RCA) * fcut(rjk, RCA);
            float cs = (ax * (float)ex + ay * (float)ey + az * (float)ez) / (r * r2k);
            float rr = r * r + r2k * r2k + rjk * rjk;
            float x0 = __expf(-0.005f * rr), x1 = __expf(-0.02f * rr);
            const float zet[4] = {1.0f, 2.0f, 4.0f, 16.0f};
            int t = 32;
            for (int zi = 0; zi < 4; ++zi)
                for (int li = 0; li < 2; ++li) {
                    float lam = li == 0 ? 1.0f : -1.0f;
                    float c = 1.0f + lam * cs;
                    float pw = zi == 0 ? c : (zi == 1 ? c * c : (zi == 2 ? (c * c) * (c * c) : 0.0f));
                    if (zi == 3) {
                        float c2 = c * c, c4 = c2 * c2, c8 = c4 * c4;
                        pw = c8 * c8;
                    }
                    float pre = exp2f(1.0f - zet[zi]) * pw * base;
                    g[t++] += pre * x0;
                    g[t++] += pre * x1;
                }
        }
    }
    const double v[3] = {qvac[3 * q], qvac[3 * q + 1], qvac[3 * q + 2]};
    const double u[3] = {qmov[3 * q], qmov[3 * q + 1], qmov[3 * q + 2]};
    const double pp[3] = {pi.x, pi.y, pi.z};
    double dv2 = 0.0, du2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        double d = pp[a] - v[a];
        d -= W.L[a] * rint(d * W.invL[a]);
        dv2 += d * d;
        double d2 = pp[a] - u[a];
        d2 -= W.L[a] * rint(d2 * W.invL[a]);
        du2 += d2 * d2;
    }
    // the vacancy site and the candidate mover as pseudo-neighbours of their
    // own kinds: 8 radial G2 terms each, with the same cutoff function
    float dv = (float)sqrt(dv2), du = (float)sqrt(du2);
    const float fv = fcut(dv, RC), fu = fcut(du, RC);
    for (int t = 0; t < 8; ++t) {
        const float s = dv - (1.5f + 5.0f * t / 7.0f), m = du - 6.0f * t / 7.0f;
        g[16 + t] = __expf(-ETA_S * s * s) * fv;
        g[24 + t] = __expf(-ETA_S * m * m) * fu;
    }
    for (int t = 0; t < 48; ++t) X[(size_t)q * 48 + t] = g[t];
}

}  // namespace

// ------------------------------------------------------------ launchers

void akmc_mlp_launch_weights(const float *p, uint16_t *w1t, uint16_t *w2t, uint16_t *w3t, uint16_t *w2r,
                             uint16_t *w3r, hipStream_t s)
{
    k_weights_bf16<<<(128 * 128 + kB - 1) / kB, kB, 0, s>>>(p, w1t, w2t, w3t, w2r, w3r);
}

// forward: X[M x 48] -> Z1,H1,Z2,H2 [M x 128], Y [M x 16] (3 used)
void akmc_mlp_launch_forward(const float *X, int M, const float *p, const uint16_t *w1t, const uint16_t *w2t,
                             const uint16_t *w3t, float *Z1, float *H1, float *Z2, float *H2, float *Y, hipStream_t s)
{
    const float *b1 = p + 48 * 128, *b2 = b1 + 128 + 128 * 128, *b3 = b2 + 128 + 128 * 3;
    dim3 g((M + 63) / 64, 8);
    k_gemm<EPI_SILU, 3><<<g, kB, 0, s>>>(X, 48, w1t, 48, M, 128, b1, nullptr, Z1, H1, 128);
    k_gemm<EPI_SILU, 8><<<g, kB, 0, s>>>(H1, 128, w2t, 128, M, 128, b2, nullptr, Z2, H2, 128);
    dim3 g3((M + 63) / 64, 1);
    k_gemm<EPI_LINEAR, 8><<<g3, kB, 0, s>>>(H2, 128, w3t, 128, M, 3, b3, nullptr, nullptr, Y, 16);
}

// loss partials and gradient of the mean-squared loss; grad laid out like p
void akmc_mlp_launch_backward(const float *X, const float *T, int M, const uint16_t *w2r, const uint16_t *w3r,
                              const float *Z1, const float *H1, const float *Z2, const float *H2, const float *Y,
                              float *dY, float *dZ2, float *dZ1, double *loss_part, float *P, int rows_per_chunk,
                              float *grad, hipStream_t s)
{
    k_mse<<<(M + kB - 1) / kB, kB, 0, s>>>(Y, T, M, dY, loss_part);
    dim3 g((M + 63) / 64, 8);
    // dZ2 = (dY W3^T) * silu'(Z2): K = 16 (padded output), B^T = W3 row-major padded
    k_gemm<EPI_DSILU, 1><<<g, kB, 0, s>>>(dY, 16, w3r, 16, M, 128, nullptr, Z2, nullptr, dZ2, 128);
    k_gemm<EPI_DSILU, 8><<<g, kB, 0, s>>>(dZ2, 128, w2r, 128, M, 128, nullptr, Z1, nullptr, dZ1, 128);
    const int nchunk = (M + rows_per_chunk - 1) / rows_per_chunk;
    float *gW1 = grad, *gb1 = gW1 + 48 * 128, *gW2 = gb1 + 128, *gb2 = gW2 + 128 * 128, *gW3 = gb2 + 128,
          *gb3 = gW3 + 128 * 3;
    // dW3 = H2^T dY (128 x 3)
    k_wgrad<<<dim3(2, 1, nchunk), kB, 0, s>>>(H2, 128, dY, 16, M, 128, 3, rows_per_chunk, P);
    k_sum_chunks<<<(128 * 3 + kB - 1) / kB, kB, 0, s>>>(P, nchunk, 128 * 3, gW3);
    // dW2 = H1^T dZ2 (128 x 128)
    k_wgrad<<<dim3(2, 8, nchunk), kB, 0, s>>>(H1, 128, dZ2, 128, M, 128, 128, rows_per_chunk, P);
    k_sum_chunks<<<(128 * 128 + kB - 1) / kB, kB, 0, s>>>(P, nchunk, 128 * 128, gW2);
    // dW1 = X^T dZ1 (48 x 128)
    k_wgrad<<<dim3(1, 8, nchunk), kB, 0, s>>>(X, 48, dZ1, 128, M, 48, 128, rows_per_chunk, P);
    k_sum_chunks<<<(48 * 128 + kB - 1) / kB, kB, 0, s>>>(P, nchunk, 48 * 128, gW1);
    k_colsum<<<3, kB, 0, s>>>(dY, 16, M, gb3);
    k_colsum<<<128, kB, 0, s>>>(dZ2, 128, M, gb2);
    k_colsum<<<128, kB, 0, s>>>(dZ1, 128, M, gb1);
}

void akmc_mlp_launch_adam(float *p, const float *g, float *m1, float *m2, int n, float lr, int t, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    k_adam<<<(n + kB - 1) / kB, kB, 0, s>>>(p, g, m1, m2, n, lr, b1, b2, 1.0e-8f, c1, c2);
}

// the training loss of one minibatch, added to a device accumulator: the
// batch's per-block partials summed in block order and scaled by 1/rows
__global__ void k_loss_accum(const double *part, int nb, double scale, double *acc)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double t = 0.0;
    for (int i = 0; i < nb; ++i) t += part[i];
    *acc += t * scale;
}

void akmc_mlp_launch_loss_accum(const double *part, int nb, double scale, double *acc, hipStream_t s)
{
    k_loss_accum<<<1, 64, 0, s>>>(part, nb, scale, acc);
}

void akmc_mlp_launch_gather(const float *src, int width, const int *idx, int M, float *dst, hipStream_t s)
{
    k_gather_rows<<<(M * width + kB - 1) / kB, kB, 0, s>>>(src, width, idx, M, dst);
}

void akmc_mlp_launch_descriptors(const BatchDev &b, int nq, const int *qatom, const double *qvac,
                                 const double *qmov, float *X, hipStream_t s)
{
    k_descriptors<<<(nq + kB - 1) / kB, kB, 0, s>>>(b, nq, qatom, qvac, qmov, X);
}
