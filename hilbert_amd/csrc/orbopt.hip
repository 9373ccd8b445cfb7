// orbopt.hip -- device kernels of the v2RDM-CASSCF orbital step.
//
// The step (orbopt_driver.cpp) keeps the density-fitted integrals of the full
// orbital space resident as dense slices B^Q (nq x nfull x nfull).  Every
// contraction is a product of row-major matrices (k_dgemm):
//   gO   (rs|tu), r..u in O          = BO^T BO             (O: core + active)
//   Z    Z[x r][Q] = sum_st d[x r s t] BO[Q][s t]
//   F    F[x][y] = sum_(r,Q) Z[x r][Q] BFO[y][r][Q] + sum_r D[x][r] h[y][occ r]
//   Jd   (pp|tu)                     = Bdiag BO
//   Kd   (ps|pu)                     = BFO[p] BFO[p]^T, batched over p
//   U    exp(K): Taylor terms and squarings
//   B^Q  U^T B^Q U, batched over Q; h the same
// and the per-rotation gradient and exact diagonal Hessian read those
// intermediates (k_oo_grad_hess).  The formulas restate the focas optimiser
// (focas/orbopt_gradient.cc, orbopt_hessian.cc) for every rotation class at
// once through the generalized Fock matrix and the second derivative of the
// energy along one rotation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "orbopt.hpp"

namespace v2h {

namespace {

constexpr int kT = 64;    // output tile
constexpr int kKc = 16;   // K chunk staged in LDS
#ifndef V2H_DGEMM_MFMA
#define V2H_DGEMM_MFMA 1
#endif
constexpr bool kUseMfma = V2H_DGEMM_MFMA != 0;  // tiled products on the matrix cores

// kc > 0: split K; batch entry bz covers k in [bz kc, min(K, (bz + 1) kc)) and
// writes its partial product to C + bz sC (summed by k_dgemm_slices).
// dg is added on the diagonal of the result (m == n).
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_dgemm(int M, int N, int K, double alpha, const double* __restrict__ A,
                                               int64_t lda, int64_t sA, const double* __restrict__ B, int64_t ldb,
                                               int64_t sB, double beta, double* __restrict__ C, int64_t ldc,
                                               int64_t sC, int kc, double dg) {
    __shared__ double As[kKc][kT + 1];
    __shared__ double Bs[kKc][kT + 1];
    const int64_t bz = blockIdx.z;
    A += bz * sA;
    B += bz * sB;
    C += bz * sC;
    const int kbeg = kc > 0 ? (int)bz * kc : 0;
    const int kend = kc > 0 ? min(K, kbeg + kc) : K;
    const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    for (int k0 = kbeg; k0 < kend; k0 += kKc) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int idx = t + 256 * e;
            int m, k;
            if (!TA) { m = idx >> 4; k = idx & 15; } else { k = idx >> 6; m = idx & 63; }
            double v = 0.0;
            if (m0 + m < M && k0 + k < kend)
                v = TA ? A[(int64_t)(k0 + k) * lda + (m0 + m)] : A[(int64_t)(m0 + m) * lda + (k0 + k)];
            As[k][m] = v;
            int n, k2;
            if (!TB) { k2 = idx >> 6; n = idx & 63; } else { n = idx >> 4; k2 = idx & 15; }
            double w = 0.0;
            if (n0 + n < N && k0 + k2 < kend)
                w = TB ? B[(int64_t)(n0 + n) * ldb + (k0 + k2)] : B[(int64_t)(k0 + k2) * ldb + (n0 + n)];
            Bs[k2][n] = w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKc; kk++) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = __fma_rn(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int m = m0 + ty + 16 * i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = n0 + tx + 16 * j;
            if (n >= N) continue;
            double* c = C + (int64_t)m * ldc + n;
            double v = beta != 0.0 ? __fma_rn(alpha, acc[i][j], beta * *c) : alpha * acc[i][j];
            if (m == n) v += dg;
            *c = v;
        }
    }
}

// The same product on the FP64 matrix cores: four wavefronts per 64 x 64
// tile, each a 32 x 32 quarter as 2 x 2 MFMA 16 x 16 x 4 tiles, operands read
// from the same LDS staging.  v_mfma_f64_16x16x4f64 lane layout (checked on
// gfx950): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and holds
// D[4 r + l / 16][l % 16] in its register r.
typedef double v4d __attribute__((ext_vector_type(4)));

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_dgemm_mfma(int M, int N, int K, double alpha, const double* __restrict__ A,
                                                    int64_t lda, int64_t sA, const double* __restrict__ B,
                                                    int64_t ldb, int64_t sB, double beta, double* __restrict__ C,
                                                    int64_t ldc, int64_t sC, int kc, double dg) {
    __shared__ double As[kKc][kT + 1];
    __shared__ double Bs[kKc][kT + 1];
    const int64_t bz = blockIdx.z;
    A += bz * sA;
    B += bz * sB;
    C += bz * sC;
    const int kbeg = kc > 0 ? (int)bz * kc : 0;
    const int kend = kc > 0 ? min(K, kbeg + kc) : K;
    const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int li = l & 15, lk = l >> 4;
    v4d acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int k0 = kbeg; k0 < kend; k0 += kKc) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int idx = t + 256 * e;
            int m, k;
            if (!TA) { m = idx >> 4; k = idx & 15; } else { k = idx >> 6; m = idx & 63; }
            double v = 0.0;
            if (m0 + m < M && k0 + k < kend)
                v = TA ? A[(int64_t)(k0 + k) * lda + (m0 + m)] : A[(int64_t)(m0 + m) * lda + (k0 + k)];
            As[k][m] = v;
            int n, k2;
            if (!TB) { k2 = idx >> 6; n = idx & 63; } else { n = idx >> 4; k2 = idx & 15; }
            double u = 0.0;
            if (n0 + n < N && k0 + k2 < kend)
                u = TB ? B[(int64_t)(n0 + n) * ldb + (k0 + k2)] : B[(int64_t)(k0 + k2) * ldb + (n0 + n)];
            Bs[k2][n] = u;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 4) {
            const double a0 = As[kk + lk][wm + li], a1 = As[kk + lk][wm + 16 + li];
            const double b0 = Bs[kk + lk][wn + li], b1 = Bs[kk + lk][wn + 16 + li];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * r + lk, n = n0 + wn + 16 * j + li;
                if (m >= M || n >= N) continue;
                double* c = C + (int64_t)m * ldc + n;
                double v = beta != 0.0 ? __fma_rn(alpha, acc[i][j][r], beta * *c) : alpha * acc[i][j][r];
                if (m == n) v += dg;
                *c = v;
            }
}

// Small products (M, N, K <= kSmallMax): a 16 x 16 output tile per
// workgroup, one output per thread, with the tile's rows of op(A) and columns
// of B staged whole in LDS (no K loop of global loads): for the chains of
// 100 x 100 products of the exponential, which are latency-bound.
constexpr int kSmallMax = 128;

template <bool TA>
__global__ __launch_bounds__(256) void k_dgemm_small(int M, int N, int K, double alpha, const double* __restrict__ A,
                                                     int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                     double* __restrict__ C, int64_t ldc, double dg) {
    __shared__ double As[16][kSmallMax + 1];
    __shared__ double Bs[kSmallMax][17];
    const int m0 = blockIdx.y * 16, n0 = blockIdx.x * 16;
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    for (int e = t; e < 16 * K; e += 256) {
        int m, k;
        if (!TA) { m = e / K; k = e - m * K; } else { k = e >> 4; m = e & 15; }
        As[m][k] = (m0 + m < M) ? (TA ? A[(int64_t)k * lda + m0 + m] : A[(int64_t)(m0 + m) * lda + k]) : 0.0;
    }
    for (int e = t; e < 16 * K; e += 256) {
        const int k = e >> 4, n = e & 15;
        Bs[k][n] = (n0 + n < N) ? B[(int64_t)k * ldb + n0 + n] : 0.0;
    }
    __syncthreads();
    double acc = 0.0;
    for (int k = 0; k < K; k++) acc = __fma_rn(As[ty][k], Bs[k][tx], acc);
    const int m = m0 + ty, n = n0 + tx;
    if (m < M && n < N) C[(int64_t)m * ldc + n] = alpha * acc + (m == n ? dg : 0.0);
}

// C[m][n] = alpha sum_s W[s][m][n] + beta C[m][n], slices summed in order
__global__ __launch_bounds__(256) void k_dgemm_slices(const double* __restrict__ W, int slices, int M, int N,
                                                      double alpha, double beta, double* __restrict__ C, int64_t ldc) {
    const int64_t mn = (int64_t)M * N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn; e += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < slices; k++) s += W[k * mn + e];
        const int m = (int)(e / N), n = (int)(e - (int64_t)m * N);
        double* c = C + (int64_t)m * ldc + n;
        *c = beta != 0.0 ? __fma_rn(alpha, s, beta * *c) : alpha * s;
    }
}

__global__ __launch_bounds__(256) void k_df_unpack(const double* __restrict__ qmo, int nq, int nfull,
                                                   double* __restrict__ bd) {
    const int64_t nn = (int64_t)nfull * nfull, total = nn * nq;
    const int64_t ntri = (int64_t)nfull * (nfull + 1) / 2;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t Q = e / nn, pq = e - Q * nn;
        const int p = (int)(pq / nfull), q = (int)(pq - (int64_t)p * nfull);
        const int64_t a = p >= q ? (int64_t)p * (p + 1) / 2 + q : (int64_t)q * (q + 1) / 2 + p;
        bd[e] = qmo[Q * ntri + a];
    }
}

__global__ __launch_bounds__(256) void k_df_pack(const double* __restrict__ bd, int nq, int nfull,
                                                 double* __restrict__ qmo) {
    const int64_t nn = (int64_t)nfull * nfull, total = nn * nq;
    const int64_t ntri = (int64_t)nfull * (nfull + 1) / 2;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t Q = e / nn, pq = e - Q * nn;
        const int p = (int)(pq / nfull), q = (int)(pq - (int64_t)p * nfull);
        if (p < q) continue;
        qmo[Q * ntri + (int64_t)p * (p + 1) / 2 + q] = bd[e];
    }
}

// spin-free D of two active (Pitzer) orbitals
__device__ double oo_d1(const OoTables& t, const double* x, int i, int j) {
    const int64_t o = t.d1[i * t.n + j];
    return o < 0 ? 0.0 : x[o] + x[t.d1[t.n * t.n + i * t.n + j]];
}

__device__ double oo_ab(const OoTables& t, const double* x, int i, int j, int k, int l) {
    const int4 a = t.pair[i * t.n + j], b = t.pair[k * t.n + l];
    if (a.x != b.x) return 0.0;
    return x[t.offab[a.x] + (int64_t)a.y * t.gab[a.x] + b.y];
}

__device__ double oo_aa(const OoTables& t, const double* x, int i, int j, int k, int l) {
    const int4 a = t.pair[i * t.n + j], b = t.pair[k * t.n + l];
    if (a.z < 0 || b.z < 0 || a.x != b.x) return 0.0;
    const int64_t e = (int64_t)a.z * t.gaa[a.x] + b.z;
    return (double)(a.w * b.w) * (x[t.offaa[a.x] + e] + x[t.offbb[a.x] + e]);
}

// d_pqrs = sum <a+_p a+_r a_s a_q> over O positions (core first, then active)
__device__ double oo_raw_d(const OoTables& t, const double* x, int p, int q, int r, int s) {
    const int nc = t.ncore;
    const bool cp = p < nc, cq = q < nc, cr = r < nc, cs = s < nc;
    const int ncnt = (int)cp + (int)cq + (int)cr + (int)cs;
    if (ncnt == 0) {
        p -= nc; q -= nc; r -= nc; s -= nc;
        return oo_ab(t, x, p, r, q, s) + oo_ab(t, x, r, p, s, q) + oo_aa(t, x, p, r, q, s);
    }
    if (ncnt == 4) return 4.0 * (p == q && r == s) - 2.0 * (p == s && q == r);
    if (ncnt != 2) return 0.0;
    if (cp && cq) return p == q ? 2.0 * oo_d1(t, x, r - nc, s - nc) : 0.0;
    if (cr && cs) return r == s ? 2.0 * oo_d1(t, x, p - nc, q - nc) : 0.0;
    if (cp && cs) return p == s ? -oo_d1(t, x, q - nc, r - nc) : 0.0;
    if (cq && cr) return q == r ? -oo_d1(t, x, p - nc, s - nc) : 0.0;
    return 0.0;
}

__global__ __launch_bounds__(256) void k_oo_density(OoTables t, const double* __restrict__ x, double* __restrict__ D,
                                                    double* __restrict__ d) {
    const int no = t.no;
    const int64_t n2 = (int64_t)no * no, n4 = n2 * n2;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n4 + n2; e += (int64_t)gridDim.x * blockDim.x) {
        if (e >= n4) {
            const int64_t f = e - n4;
            const int a = (int)(f / no), b = (int)(f - (int64_t)a * no);
            double v;
            if (a < t.ncore || b < t.ncore) v = (a == b) ? 2.0 : 0.0;
            else v = oo_d1(t, x, a - t.ncore, b - t.ncore);
            if ((a < t.ncore) != (b < t.ncore)) v = 0.0;
            D[f] = v;
            continue;
        }
        const int p = (int)(e / (n2 * no));
        const int q = (int)((e / n2) % no);
        const int r = (int)((e / no) % no);
        const int s = (int)(e % no);
        d[e] = 0.25 * (oo_raw_d(t, x, p, q, r, s) + oo_raw_d(t, x, q, p, r, s) + oo_raw_d(t, x, p, q, s, r) +
                       oo_raw_d(t, x, q, p, s, r));
    }
}

__global__ __launch_bounds__(256) void k_oo_gather(const double* __restrict__ bd, int nq, int nfull,
                                                   const int* __restrict__ occ, int no, double* __restrict__ BO,
                                                   double* __restrict__ BFO, double* __restrict__ Bdiag) {
    const int64_t nn = (int64_t)nfull * nfull;
    const int64_t n1 = (int64_t)nq * no * no, n2 = (int64_t)nfull * no * nq, n3 = (int64_t)nfull * nq;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n1 + n2 + n3; e += (int64_t)gridDim.x * blockDim.x) {
        if (e < n1) {
            const int64_t Q = e / ((int64_t)no * no);
            const int rs = (int)(e - Q * no * no), r = rs / no, s = rs - r * no;
            BO[e] = bd[Q * nn + (int64_t)occ[r] * nfull + occ[s]];
        } else if (e < n1 + n2) {
            const int64_t f = e - n1;
            const int64_t p = f / ((int64_t)no * nq);
            const int64_t rQ = f - p * no * nq;
            const int r = (int)(rQ / nq), Q = (int)(rQ - (int64_t)r * nq);
            BFO[f] = bd[(int64_t)Q * nn + p * nfull + occ[r]];
        } else {
            const int64_t f = e - n1 - n2;
            const int64_t p = f / nq, Q = f - p * nq;
            Bdiag[f] = bd[Q * nn + p * nfull + p];
        }
    }
}

__global__ __launch_bounds__(256) void k_oo_fock1(const double* __restrict__ D, const double* __restrict__ h,
                                                  const int* __restrict__ occ, int no, int nfull, double* __restrict__ F) {
    const int64_t total = (int64_t)no * nfull;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int xo = (int)(e / nfull), y = (int)(e - (int64_t)xo * nfull);
        double acc = 0.0;
        for (int r = 0; r < no; r++) acc = __fma_rn(D[xo * no + r], h[(int64_t)y * nfull + occ[r]], acc);
        F[e] += acc;
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One wavefront per rotation pair; the lanes split the O^2 sums and combine
// them with butterfly shuffles (the same order on every run).
__global__ __launch_bounds__(256) void k_oo_grad_hess(OoGradArgs a) {
    const int k = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= a.nrot) return;  // uniform per wavefront
    const int p = a.pairs[k].x, q = a.pairs[k].y;
    const int no = a.no, nf = a.nfull;
    const int n2 = no * no;
    const int ip = a.pos[p], iq = a.pos[q];
    const double* Jp = a.Jd + (int64_t)p * n2;
    const double* Jq = a.Jd + (int64_t)q * n2;
    const double* Kp = a.Kd + (int64_t)p * n2;
    const double* Kq = a.Kd + (int64_t)q * n2;
    const int64_t no3 = (int64_t)n2 * no;
    double g12 = 0.0, g13 = 0.0, c12 = 0.0, c13 = 0.0;
    for (int e = lane; e < n2; e += 64) {
        const int s = e / no, u = e - s * no;
        if (ip >= 0) {
            g12 = __fma_rn(a.d[((int64_t)ip * no + ip) * n2 + e], Jq[e], g12);
            g13 = __fma_rn(a.d[ip * no3 + (int64_t)s * n2 + (int64_t)ip * no + u], Kq[e], g13);
        }
        if (iq >= 0) {
            g12 = __fma_rn(a.d[((int64_t)iq * no + iq) * n2 + e], Jp[e], g12);
            g13 = __fma_rn(a.d[iq * no3 + (int64_t)s * n2 + (int64_t)iq * no + u], Kp[e], g13);
        }
        if (ip >= 0 && iq >= 0) {
            const int64_t row = ((int64_t)ip * no + iq) * n2;
            c12 = __fma_rn(a.d[row + e], a.gO[row + e], c12);
            c13 = __fma_rn(a.d[ip * no3 + (int64_t)s * n2 + (int64_t)iq * no + u],
                           a.gO[iq * no3 + (int64_t)s * n2 + (int64_t)ip * no + u], c13);
        }
    }
    g12 = wave_sum(g12);
    g13 = wave_sum(g13);
    c12 = wave_sum(c12);
    c13 = wave_sum(c13);
    if (lane != 0) return;
    const double Fpq = ip >= 0 ? a.F[(int64_t)ip * nf + q] : 0.0;
    const double Fqp = iq >= 0 ? a.F[(int64_t)iq * nf + p] : 0.0;
    a.g[k] = 2.0 * (Fqp - Fpq);
    const double Fpp = ip >= 0 ? a.F[(int64_t)ip * nf + p] : 0.0;
    const double Fqq = iq >= 0 ? a.F[(int64_t)iq * nf + q] : 0.0;
    const double Dpp = ip >= 0 ? a.D[ip * no + ip] : 0.0;
    const double Dqq = iq >= 0 ? a.D[iq * no + iq] : 0.0;
    const double Dpq = (ip >= 0 && iq >= 0) ? a.D[ip * no + iq] : 0.0;
    const double hpp = a.h[(int64_t)p * nf + p], hqq = a.h[(int64_t)q * nf + q], hpq = a.h[(int64_t)p * nf + q];
    const double val = -2.0 * (Fpp + Fqq) + 2.0 * (Dpp * hqq + Dqq * hpp - 2.0 * Dpq * hpq);
    a.H[k] = val + 2.0 * (g12 - 2.0 * c12) + 4.0 * (g13 - 2.0 * c13);
}

constexpr int kRed = 1024;

__device__ void block_sum2(double& a, double& b, double* sa, double* sb) {
    const int t = threadIdx.x;
    sa[t] = a;
    sb[t] = b;
    __syncthreads();
    for (int w = kRed / 2; w > 0; w >>= 1) {
        if (t < w) {
            sa[t] += sa[t + w];
            sb[t] += sb[t + w];
        }
        __syncthreads();
    }
    a = sa[0];
    b = sb[0];
}

__global__ __launch_bounds__(kRed) void k_oo_precondition(const double* __restrict__ g, const double* __restrict__ H,
                                                          int nrot, double r, double* __restrict__ kref,
                                                          double* __restrict__ kcur, double* __restrict__ ksave,
                                                          double* __restrict__ out) {
    __shared__ double sa[kRed], sb[kRed];
    double gg = 0.0, quad = 0.0;
    for (int k = threadIdx.x; k < nrot; k += kRed) {
        const double gk = g[k], hk = H[k], ck = kcur[k];
        gg = __fma_rn(gk, gk, gg);
        quad += gk * ck + 0.5 * ck * hk * ck;
        double hv = hk < 0.0 ? -hk : hk;
        if (hv < 1.0e-2) hv = 1.0e-2;
        const double kr = -gk / hv;
        kref[k] = kr;
        kcur[k] = r * kr;
        ksave[k] = r * kr;
    }
    block_sum2(gg, quad, sa, sb);
    if (threadIdx.x == 0) {
        out[0] = gg;
        out[1] = quad;
    }
}

__global__ __launch_bounds__(256) void k_oo_update_kappa(const double* __restrict__ kref, int nrot, double r,
                                                         double* __restrict__ kcur, double* __restrict__ ksave) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrot) return;
    const double kn = r * kref[k] - kcur[k];
    kcur[k] = kn;
    ksave[k] += kn;
}

__global__ __launch_bounds__(256) void k_oo_fill(double* __restrict__ p, int64_t n, double v) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) p[e] = v;
}

__global__ __launch_bounds__(256) void k_oo_scatter_kappa(const double* __restrict__ kappa, const int2* __restrict__ pairs,
                                                          int nrot, double scale, int nfull, double* __restrict__ K) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrot) return;
    const int p = pairs[k].x, q = pairs[k].y;
    const double v = scale * kappa[k];
    K[(int64_t)p * nfull + q] = v;
    K[(int64_t)q * nfull + p] = -v;
}

__global__ __launch_bounds__(kRed) void k_oo_norm1(const double* __restrict__ K, int nfull, double* __restrict__ out) {
    __shared__ double sa[kRed], sb[kRed];
    double mx = 0.0, unused = 0.0;
    for (int j = threadIdx.x; j < nfull; j += kRed) {
        double s = 0.0;
        for (int i = 0; i < nfull; i++) s += fabs(K[(int64_t)i * nfull + j]);
        mx = s > mx ? s : mx;
    }
    sa[threadIdx.x] = mx;
    sb[threadIdx.x] = unused;
    __syncthreads();
    for (int w = kRed / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sa[0];
}

__global__ __launch_bounds__(256) void k_oo_identity(double* __restrict__ U, int nfull) {
    const int64_t n = (int64_t)nfull * nfull;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        U[e] = (e / nfull == e % nfull) ? 1.0 : 0.0;
}

__global__ __launch_bounds__(256) void k_oo_axpy(double a, const double* x, double* y, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        y[e] = __fma_rn(a, x[e], y[e]);
}

__global__ __launch_bounds__(256) void k_oo_sub(const double* __restrict__ x, const double* __restrict__ z,
                                                double* __restrict__ y, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        y[e] = x[e] - z[e];
}

unsigned grid_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

}  // namespace

namespace {

void dgemm_raw(bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int64_t lda, int64_t sA,
               const double* B, int64_t ldb, int64_t sB, double beta, double* C, int64_t ldc, int64_t sC, int batch,
               int kc, double dg, hipStream_t s) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    const int64_t gy = (M + kT - 1) / kT, gx = (N + kT - 1) / kT;
    if (gy > 65535 || batch > 65535 || gx > (int64_t)1 << 30)
        throw std::invalid_argument("launch_dgemm: shape " + std::to_string(M) + " x " + std::to_string(N) +
                                    " x batch " + std::to_string(batch) + " exceeds the grid");
    const dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)batch);
    if (kUseMfma) {
        if (!ta && !tb) hipLaunchKernelGGL((k_dgemm_mfma<false, false>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, kc, dg);
        else if (ta && !tb) hipLaunchKernelGGL((k_dgemm_mfma<true, false>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, kc, dg);
        else if (!ta && tb) hipLaunchKernelGGL((k_dgemm_mfma<false, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, kc, dg);
        else hipLaunchKernelGGL((k_dgemm_mfma<true, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, kc, dg);
        return;
    }
    if (!ta && !tb) hipLaunchKernelGGL((k_dgemm<false, false>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, kc, dg);
    else if (ta && !tb) hipLaunchKernelGGL((k_dgemm<true, false>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, kc, dg);
    else if (!ta && tb) hipLaunchKernelGGL((k_dgemm<false, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, kc, dg);
    else hipLaunchKernelGGL((k_dgemm<true, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, kc, dg);
}

}  // namespace

void launch_dgemm(bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int64_t lda, int64_t sA,
                  const double* B, int64_t ldb, int64_t sB, double beta, double* C, int64_t ldc, int64_t sC,
                  int batch, hipStream_t s, double dg) {
    dgemm_raw(ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, 0, dg, s);
}

int dgemm_split_slices(int M, int N, int K) {
    const int64_t tiles = (int64_t)((M + kT - 1) / kT) * ((N + kT - 1) / kT);
    int sl = (int)std::min<int64_t>(kMaxSlices, std::max<int64_t>(1, 512 / std::max<int64_t>(1, tiles)));
    sl = std::min(sl, std::max(1, K / 64));
    return sl;
}

void launch_dgemm_small(bool ta, int M, int N, int K, double alpha, const double* A, int64_t lda, const double* B,
                        int64_t ldb, double* C, int64_t ldc, hipStream_t s, double dg) {
    if (M <= 0 || N <= 0) return;
    if (M > kSmallMax || N > kSmallMax || K > kSmallMax || K < 1) {
        launch_dgemm(ta, false, M, N, K, alpha, A, lda, 0, B, ldb, 0, 0.0, C, ldc, 0, 1, s, dg);
        return;
    }
    const dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16));
    if (ta) hipLaunchKernelGGL((k_dgemm_small<true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, C, ldc, dg);
    else hipLaunchKernelGGL((k_dgemm_small<false>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, C, ldc, dg);
}

void launch_dgemm_split(bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* work,
                        hipStream_t s) {
    const int sl = dgemm_split_slices(M, N, K);
    if (sl <= 1) {
        dgemm_raw(ta, tb, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1, 0, 0.0, s);
        return;
    }
    const int kc = ((K + sl - 1) / sl + kKc - 1) / kKc * kKc;
    const int slices = (K + kc - 1) / kc;
    dgemm_raw(ta, tb, M, N, K, 1.0, A, lda, 0, B, ldb, 0, 0.0, work, N, (int64_t)M * N, slices, kc, 0.0, s);
    const int64_t mn = (int64_t)M * N;
    hipLaunchKernelGGL(k_dgemm_slices, dim3(grid_for(mn)), dim3(256), 0, s, work, slices, M, N, alpha, beta, C, ldc);
}

void launch_df_unpack(const double* qmo, int nq, int nfull, double* bd, hipStream_t s) {
    hipLaunchKernelGGL(k_df_unpack, dim3(grid_for((int64_t)nq * nfull * nfull)), dim3(256), 0, s, qmo, nq, nfull, bd);
}

void launch_df_pack(const double* bd, int nq, int nfull, double* qmo, hipStream_t s) {
    hipLaunchKernelGGL(k_df_pack, dim3(grid_for((int64_t)nq * nfull * nfull)), dim3(256), 0, s, bd, nq, nfull, qmo);
}

void launch_oo_density(const OoTables& t, const double* x, double* D, double* d, hipStream_t s) {
    const int64_t n2 = (int64_t)t.no * t.no;
    hipLaunchKernelGGL(k_oo_density, dim3(grid_for(n2 * n2 + n2)), dim3(256), 0, s, t, x, D, d);
}

void launch_oo_gather(const double* bd, int nq, int nfull, const int* occ, int no, double* BO, double* BFO,
                      double* Bdiag, hipStream_t s) {
    const int64_t total = (int64_t)nq * no * no + (int64_t)nfull * no * nq + (int64_t)nfull * nq;
    hipLaunchKernelGGL(k_oo_gather, dim3(grid_for(total)), dim3(256), 0, s, bd, nq, nfull, occ, no, BO, BFO, Bdiag);
}

void launch_oo_fock1(const double* D, const double* h, const int* occ, int no, int nfull, double* F, hipStream_t s) {
    hipLaunchKernelGGL(k_oo_fock1, dim3(grid_for((int64_t)no * nfull)), dim3(256), 0, s, D, h, occ, no, nfull, F);
}

void launch_oo_grad_hess(const OoGradArgs& a, hipStream_t s) {
    if (a.nrot <= 0) return;
    hipLaunchKernelGGL(k_oo_grad_hess, dim3((a.nrot + 3) / 4), dim3(256), 0, s, a);
}

void launch_oo_precondition(const double* g, const double* H, int nrot, double r, double* kref, double* kcur,
                            double* ksave, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_oo_precondition, dim3(1), dim3(kRed), 0, s, g, H, nrot, r, kref, kcur, ksave, out);
}

void launch_oo_update_kappa(const double* kref, int nrot, double r, double* kcur, double* ksave, hipStream_t s) {
    if (nrot <= 0) return;
    hipLaunchKernelGGL(k_oo_update_kappa, dim3((nrot + 255) / 256), dim3(256), 0, s, kref, nrot, r, kcur, ksave);
}

void launch_oo_kappa_matrix(const double* kappa, const int2* pairs, int nrot, double scale, int nfull, double* K,
                            double* out, hipStream_t s) {
    const int64_t nn = (int64_t)nfull * nfull;
    hipLaunchKernelGGL(k_oo_fill, dim3(grid_for(nn)), dim3(256), 0, s, K, nn, 0.0);
    if (nrot > 0)
        hipLaunchKernelGGL(k_oo_scatter_kappa, dim3((nrot + 255) / 256), dim3(256), 0, s, kappa, pairs, nrot, scale,
                           nfull, K);
    hipLaunchKernelGGL(k_oo_norm1, dim3(1), dim3(kRed), 0, s, K, nfull, out);
}

void launch_oo_identity(double* U, int nfull, hipStream_t s) {
    hipLaunchKernelGGL(k_oo_identity, dim3(grid_for((int64_t)nfull * nfull)), dim3(256), 0, s, U, nfull);
}

void launch_oo_axpy(double a, const double* x, double* y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_oo_axpy, dim3(grid_for(n)), dim3(256), 0, s, a, x, y, n);
}

void launch_oo_sub(const double* x, const double* z, double* y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_oo_sub, dim3(grid_for(n)), dim3(256), 0, s, x, z, y, n);
}

}  // namespace v2h
