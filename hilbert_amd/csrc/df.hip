// df.hip -- density-fitted Hamiltonian intake on the device: the active-space
// integrals of one instance from three-index (Q|mn) integrals, with the
// frozen-core terms.
//
// The reference forms every integral it packs by a strided dot over the
// auxiliary index (the DF branch of TEI, tei.cc:280-283), called from
// RepackIntegrals (tei.cc:139-199) and FrozenCoreEnergy (tei.cc:201-275).
// Here the same sums are products of pair-major row blocks:
//   rows   R[r][Q] = qmo[Q ntri + INDEX(i_r, j_r)] for a list of orbital pairs
//   g      G = Pa Pa^T        Pa: rows of the active pairs (p, q), p, q < n
//   J      d_Q = sum_k R(k,k)[Q],  J_pq = Pa[pq] . d
//   K      K = M M^T          M[p]: rows (a_p, k) for every core k, contiguous
//   h'     h_pq + (2 J_pq - K_pq)
//   E_fzc  2 sum_k h_kk + 2 d.d - sum_(k,l) |R(k,l)|^2
// G lands directly in the (pq|rs) part of the instance's integral source
// array and h' in its one-electron part, the layout the cost-vector gather
// (k_pack_cost) reads, so the packing is the same as for host integrals.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "df.hpp"

namespace v2h {

namespace {

constexpr int kT = 64;   // output tile of the row-product kernel
constexpr int kK = 16;   // auxiliary-index chunk staged in LDS

// R[r][Q] = qmo[Q ntri + idx[r]]; Q fastest, so the stores coalesce and the
// loads of one Q row stay inside its ntri-long slice.
__global__ __launch_bounds__(256) void k_df_rows(const double* qmo, int64_t ntri, int nq, const int64_t* idx,
                                                 int64_t m, double* R) {
    const int64_t total = m * nq;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / nq;
        const int q = (int)(e - r * nq);
        R[e] = qmo[(int64_t)q * ntri + idx[r]];
    }
}

// C[a][b] = sum_k X[a][k] Y[b][k] for a < ma, b < mb, k < kk: 64 x 64 output
// tiles, 4 x 4 per thread, k in LDS-staged chunks of 16.  Every output is one
// sequential fused multiply-add chain over k.
__global__ __launch_bounds__(256) void k_df_gemm_nt(const double* X, int64_t ldx, const double* Y, int64_t ldy,
                                                    int ma, int mb, int64_t kk, double* C, int64_t ldc) {
    __shared__ double xs[kK][kT + 1], ys[kK][kT + 1];
    const int a0 = blockIdx.y * kT, b0 = blockIdx.x * kT;
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    for (int64_t q0 = 0; q0 < kk; q0 += kK) {
        for (int e = t; e < kT * kK; e += 256) {
            const int r = e / kK, k = e % kK;
            const int64_t q = q0 + k;
            xs[k][r] = (a0 + r < ma && q < kk) ? X[(int64_t)(a0 + r) * ldx + q] : 0.0;
            ys[k][r] = (b0 + r < mb && q < kk) ? Y[(int64_t)(b0 + r) * ldy + q] : 0.0;
        }
        __syncthreads();
        const int kn = (int)((kk - q0) < kK ? (kk - q0) : kK);
        for (int k = 0; k < kn; k++) {
            double xv[4], yv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                xv[i] = xs[k][ty + 16 * i];
                yv[i] = ys[k][tx + 16 * i];
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = fma(xv[i], yv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int a = a0 + ty + 16 * i, b = b0 + tx + 16 * j;
            if (a < ma && b < mb) C[(int64_t)a * ldc + b] = acc[i][j];
        }
}

// d[Q] = sum over the core orbitals k, in order, of R(k,k)[Q]
__global__ __launch_bounds__(256) void k_df_dvec(const double* Rkk, int ncore, int nq, double* d) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < ncore; c++) s += Rkk[(int64_t)c * nq + q];
        d[q] = s;
    }
}

__device__ __forceinline__ double block_sum256(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double s = ((red[0] + red[1]) + red[2]) + red[3];
    __syncthreads();
    return s;
}

// One workgroup: h'[pq] = h_up(a_p, a_q) + (2 Pa[pq] . d - K[pq]) and E_fzc.
__global__ __launch_bounds__(256) void k_df_finish(const double* h_full, int nfull, const int* active, int n,
                                                   const int* core, int ncore, const double* Pa, const double* d,
                                                   const double* Kc, const double* Rcc, int nq, double* h_out,
                                                   double* efzc) {
    __shared__ double red[4];
    const int t = threadIdx.x;
    for (int pq = t; pq < n * n; pq += 256) {
        const int i = active[pq / n], j = active[pq % n];
        const double h = i <= j ? h_full[(int64_t)i * nfull + j] : h_full[(int64_t)j * nfull + i];
        double jv = 0.0;
        if (ncore > 0)
            for (int q = 0; q < nq; q++) jv = fma(Pa[(int64_t)pq * nq + q], d[q], jv);
        h_out[pq] = ncore > 0 ? h + (2.0 * jv - Kc[pq]) : h;
    }
    // E_fzc = 2 sum_k h_kk + 2 d.d - sum_(k,l) |R(k,l)|^2
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int c = t; c < ncore; c += 256) s1 += h_full[(int64_t)core[c] * nfull + core[c]];
    for (int q = t; q < nq && ncore > 0; q += 256) s2 = fma(d[q], d[q], s2);
    const int64_t nrc = (int64_t)ncore * ncore * nq;
    for (int64_t e = t; e < nrc; e += 256) s3 = fma(Rcc[e], Rcc[e], s3);
    s1 = block_sum256(s1, red);
    s2 = block_sum256(s2, red);
    s3 = block_sum256(s3, red);
    if (t == 0) *efzc = 2.0 * s1 + (2.0 * s2 - s3);
}

int grid_of(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (int)(g < 2048 ? (g > 0 ? g : 1) : 2048);
}

}  // namespace

void launch_df_intake(const DfArgs& a, hipStream_t s) {
    const int64_t m = a.m_rows;
    hipLaunchKernelGGL(k_df_rows, dim3(grid_of(m * a.nq)), dim3(256), 0, s, a.qmo, a.ntri, a.nq, a.idx, m, a.rows);
    const int n = a.n, n2 = n * n;
    const double* Pa = a.rows;  // rows [0, n^2): active pairs
    // (pq|rs) = Pa Pa^T straight into the source array, row pq = p n + q
    hipLaunchKernelGGL(k_df_gemm_nt, dim3((n2 + kT - 1) / kT, (n2 + kT - 1) / kT), dim3(256), 0, s, Pa,
                       (int64_t)a.nq, Pa, (int64_t)a.nq, n2, n2, (int64_t)a.nq, a.g_out, (int64_t)n2);
    const double* Rkk = a.rows + (int64_t)n2 * a.nq;                 // ncore rows (k, k)
    const double* M = Rkk + (int64_t)a.ncore * a.nq;                 // n x ncore rows (a_p, k)
    const double* Rcc = M + (int64_t)n * a.ncore * a.nq;             // ncore^2 rows (k, l)
    if (a.ncore > 0) {
        hipLaunchKernelGGL(k_df_dvec, dim3(grid_of(a.nq)), dim3(256), 0, s, Rkk, a.ncore, a.nq, a.dvec);
        const int64_t kk = (int64_t)a.ncore * a.nq;
        hipLaunchKernelGGL(k_df_gemm_nt, dim3((n + kT - 1) / kT, (n + kT - 1) / kT), dim3(256), 0, s, M, kk, M, kk,
                           n, n, kk, a.kmat, (int64_t)n);
    }
    hipLaunchKernelGGL(k_df_finish, dim3(1), dim3(256), 0, s, a.h_full, a.nfull, a.active, n, a.core, a.ncore, Pa,
                       a.dvec, a.kmat, Rcc, a.nq, a.h_out, a.efzc);
}

}  // namespace v2h
