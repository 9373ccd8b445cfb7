// df.hip -- density-fitted Hamiltonian intake on the device: the active-space
// integrals of one instance from three-index (Q|mn) integrals, with the
// frozen-core terms.
//
// The reference forms every integral it packs by a strided dot over the
// auxiliary index (the DF branch of TEI, tei.cc:280-283), called from
// RepackIntegrals (tei.cc:139-199) and FrozenCoreEnergy (tei.cc:201-275).
// Here the same sums are products of pair-major row blocks:
//   rows   R[r][Q] = qmo[Q ntri + INDEX(i_r, j_r)] for a list of orbital pairs
//   g      Gu = Pu Pu^T       Pu: rows of the active pairs p >= q, row
//                             p (p + 1) / 2 + q (INDEX(a_p, a_q) does not
//                             depend on the order of p and q)
//          (pq|rs) = Gu[u(p,q)][u(r,s)], expanded to all n^4
//   J      d_Q = sum_k R(k,k)[Q],  J_pq = Pu[u(p,q)] . d
//   K      K_pq = M[p] . M[q]     M[p]: rows (a_p, k) for every core k, contiguous
//   h'     h_pq + (2 J_pq - K_pq)
//   E_fzc  2 sum_k h_kk + 2 d.d - sum_(k,l) |R(k,l)|^2
// (pq|rs) lands in the two-electron part of the instance's integral source
// array and h' in its one-electron part, the layout the cost-vector gather
// (k_pack_cost) reads, so the packing is the same as for host integrals.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "df.hpp"

namespace v2h {

namespace {

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

// C_z[a][b] = sum_{k in split z} X[a][k] Y[b][k] for a < ma, b < mb: (16 TM)^2
// output tiles, TM x TM per thread of a 16 x 16 workgroup, k in LDS-staged
// chunks of 16; blockIdx.z = z picks the k range [z kc, (z + 1) kc) and the
// output plane C + z cplane.  Every output of a plane is one sequential fused
// multiply-add chain over its k range, whatever TM.
template <int TM>
__global__ __launch_bounds__(256) void k_df_gemm_nt(const double* X, int64_t ldx, const double* Y, int64_t ldy,
                                                    int ma, int mb, int64_t kk, int64_t kc, double* C, int64_t ldc,
                                                    int64_t cplane) {
    constexpr int kT = 16 * TM;
    __shared__ double xs[kK][kT + 1], ys[kK][kT + 1];
    const int a0 = blockIdx.y * kT, b0 = blockIdx.x * kT;
    const int64_t klo = (int64_t)blockIdx.z * kc, khi = klo + kc < kk ? klo + kc : kk;
    C += (int64_t)blockIdx.z * cplane;
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    double acc[TM][TM];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TM; j++) acc[i][j] = 0.0;
    for (int64_t q0 = klo; q0 < khi; q0 += kK) {
        for (int e = t; e < kT * kK; e += 256) {
            const int r = e / kK, k = e % kK;
            const int64_t q = q0 + k;
            xs[k][r] = (a0 + r < ma && q < khi) ? X[(int64_t)(a0 + r) * ldx + q] : 0.0;
            ys[k][r] = (b0 + r < mb && q < khi) ? Y[(int64_t)(b0 + r) * ldy + q] : 0.0;
        }
        __syncthreads();
        const int kn = (int)((khi - q0) < kK ? (khi - q0) : kK);
        for (int k = 0; k < kn; k++) {
            double xv[TM], yv[TM];
#pragma unroll
            for (int i = 0; i < TM; i++) {
                xv[i] = xs[k][ty + 16 * i];
                yv[i] = ys[k][tx + 16 * i];
            }
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TM; j++) acc[i][j] = fma(xv[i], yv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TM; j++) {
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

__device__ __forceinline__ int upair(int p, int q) { return p >= q ? p * (p + 1) / 2 + q : q * (q + 1) / 2 + p; }

// (pq|rs) = sum over the planes z, in order, of Gu_z[u(p,q)][u(r,s)], for
// all p, q, r, s < n
__global__ __launch_bounds__(256) void k_df_expand(const double* Gu, int nu, int planes, int n, double* g) {
    const int64_t n2 = (int64_t)n * n, total = n2 * n2, plane = (int64_t)nu * nu;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int pq = (int)(e / n2), rs = (int)(e - (int64_t)pq * n2);
        const int64_t u = (int64_t)upair(pq / n, pq % n) * nu + upair(rs / n, rs % n);
        double v = Gu[u];
        for (int z = 1; z < planes; z++) v += Gu[z * plane + u];
        g[e] = v;
    }
}

// One workgroup per active pair (p, q): h'_pq = h_up(a_p, a_q) + (2 J - K)
// with J = Pu[u(p,q)] . d and K = M[p] . M[q] (fixed-order block sums).
__global__ __launch_bounds__(256) void k_df_fock(const double* h_full, int nfull, const int* active, int n, int ncore,
                                                 const double* Pu, const double* d, const double* M, int nq,
                                                 double* h_out) {
    __shared__ double red[4];
    const int pq = blockIdx.x, p = pq / n, q = pq % n;
    const int i = active[p], j = active[q];
    const double h = i <= j ? h_full[(int64_t)i * nfull + j] : h_full[(int64_t)j * nfull + i];
    if (ncore == 0) {
        if (threadIdx.x == 0) h_out[pq] = h;
        return;
    }
    const double* pu = Pu + (int64_t)upair(p, q) * nq;
    double jv = 0.0;
    for (int Q = threadIdx.x; Q < nq; Q += 256) jv = fma(pu[Q], d[Q], jv);
    const int64_t kk = (int64_t)ncore * nq;
    const double *mp = M + p * kk, *mq = M + q * kk;
    double kv = 0.0;
    for (int64_t e = threadIdx.x; e < kk; e += 256) kv = fma(mp[e], mq[e], kv);
    jv = block_sum256(jv, red);
    kv = block_sum256(kv, red);
    if (threadIdx.x == 0) h_out[pq] = h + (2.0 * jv - kv);
}

// One workgroup per core pair row (k, l): sq[row] = |R(k,l)|^2
__global__ __launch_bounds__(256) void k_df_rownorm(const double* Rcc, int nq, double* sq) {
    __shared__ double red[4];
    const double* r = Rcc + (int64_t)blockIdx.x * nq;
    double s = 0.0;
    for (int q = threadIdx.x; q < nq; q += 256) s = fma(r[q], r[q], s);
    s = block_sum256(s, red);
    if (threadIdx.x == 0) sq[blockIdx.x] = s;
}

// One workgroup: E_fzc = 2 sum_k h_kk + 2 d.d - sum_(k,l) sq[(k,l)]
__global__ __launch_bounds__(256) void k_df_efzc(const double* h_full, int nfull, const int* core, int ncore,
                                                 const double* d, const double* sq, int nq, double* efzc) {
    __shared__ double red[4];
    const int t = threadIdx.x;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int c = t; c < ncore; c += 256) s1 += h_full[(int64_t)core[c] * nfull + core[c]];
    for (int q = t; q < nq && ncore > 0; q += 256) s2 = fma(d[q], d[q], s2);
    for (int r = t; r < ncore * ncore; r += 256) s3 += sq[r];
    s1 = block_sum256(s1, red);
    s2 = block_sum256(s2, red);
    s3 = block_sum256(s3, red);
    if (t == 0) *efzc = 2.0 * s1 + (2.0 * s2 - s3);
}

// Tiling of the pair-product Gu = Pu Pu^T (nu x nu, k = nq): the widest
// per-thread tile that still gives at least 256 workgroups (one per CU), and
// below that, splits of k (each at least 32 long) up to 256 workgroups.
struct GemmPlan {
    int tm, splits;
    int64_t kc;
};
GemmPlan gemm_plan(int nu, int64_t kk) {
    auto tiles = [&](int tm) { return (int64_t)((nu + 16 * tm - 1) / (16 * tm)) * ((nu + 16 * tm - 1) / (16 * tm)); };
    GemmPlan p{};
    p.tm = tiles(4) >= 256 ? 4 : (tiles(2) >= 256 ? 2 : 1);
    const int64_t want = (256 + tiles(p.tm) - 1) / tiles(p.tm), most = kk / 32 > 1 ? kk / 32 : 1;
    p.splits = (int)(want < most ? want : most);
    if (p.splits > kDfMaxSplits) p.splits = kDfMaxSplits;
    p.kc = (kk + p.splits - 1) / p.splits;
    p.splits = (int)((kk + p.kc - 1) / p.kc);  // no empty split
    return p;
}

void launch_gemm_nt(const double* X, int64_t ldx, int m, int64_t kk, double* C, hipStream_t s) {
    const GemmPlan p = gemm_plan(m, kk);
    const int tw = 16 * p.tm;
    const dim3 grid((m + tw - 1) / tw, (m + tw - 1) / tw, p.splits);
    const int64_t plane = (int64_t)m * m;
    if (p.tm == 4)
        hipLaunchKernelGGL(k_df_gemm_nt<4>, grid, dim3(256), 0, s, X, ldx, X, ldx, m, m, kk, p.kc, C, (int64_t)m, plane);
    else if (p.tm == 2)
        hipLaunchKernelGGL(k_df_gemm_nt<2>, grid, dim3(256), 0, s, X, ldx, X, ldx, m, m, kk, p.kc, C, (int64_t)m, plane);
    else
        hipLaunchKernelGGL(k_df_gemm_nt<1>, grid, dim3(256), 0, s, X, ldx, X, ldx, m, m, kk, p.kc, C, (int64_t)m, plane);
}

// G[a][b] = sum_Q qmo[Q][a] qmo[Q][b] over the pair index of every orbital
// pair (the full two-electron integral table of the FCIDUMP writer): one
// workgroup per 64 x 64 tile on or above the diagonal (tile rows ta <= tile
// columns tb), 4 x 4 outputs per thread, Q staged 16 rows at a time with the
// loads running along the pair index (coalesced).  Each output is one
// sequential fused multiply-add chain over Q.
__global__ __launch_bounds__(256) void k_pair_gram(const double* qmo, int64_t ntri, int nq, double* G) {
    __shared__ double xs[kK][64], ys[kK][64];
    const int64_t k = blockIdx.x;
    int64_t tb = (int64_t)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while (tb * (tb + 1) / 2 > k) tb--;
    while ((tb + 1) * (tb + 2) / 2 <= k) tb++;
    const int64_t ta = k - tb * (tb + 1) / 2;
    const int64_t a0 = ta * 64, b0 = tb * 64;
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    for (int q0 = 0; q0 < nq; q0 += kK) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int e = t + 256 * i, r = e & 63, kq = e >> 6;
            const int q = q0 + kq;
            xs[kq][r] = (a0 + r < ntri && q < nq) ? qmo[(int64_t)q * ntri + a0 + r] : 0.0;
            ys[kq][r] = (b0 + r < ntri && q < nq) ? qmo[(int64_t)q * ntri + b0 + r] : 0.0;
        }
        __syncthreads();
        const int kn = nq - q0 < kK ? nq - q0 : kK;
        for (int kq = 0; kq < kn; kq++) {
            double xv[4], yv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                xv[i] = xs[kq][ty + 16 * i];
                yv[i] = ys[kq][tx + 16 * i];
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
            const int64_t a = a0 + ty + 16 * i, b = b0 + tx + 16 * j;
            if (a < ntri && b < ntri) G[a * ntri + b] = acc[i][j];
        }
}

int grid_of(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (int)(g < 2048 ? (g > 0 ? g : 1) : 2048);
}

}  // namespace

void launch_pair_gram(const double* qmo, int64_t ntri, int nq, double* G, hipStream_t s) {
    const int64_t T = (ntri + 63) / 64;
    hipLaunchKernelGGL(k_pair_gram, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, qmo, ntri, nq, G);
}

void launch_df_intake(const DfArgs& a, hipStream_t s) {
    const int64_t m = a.m_rows;
    hipLaunchKernelGGL(k_df_rows, dim3(grid_of(m * a.nq)), dim3(256), 0, s, a.qmo, a.ntri, a.nq, a.idx, m, a.rows);
    const int n = a.n, nu = n * (n + 1) / 2;
    const double* Pu = a.rows;                                       // nu rows (a_p, a_q), p >= q
    const double* Rkk = Pu + (int64_t)nu * a.nq;                     // ncore rows (k, k)
    const double* M = Rkk + (int64_t)a.ncore * a.nq;                 // n x ncore rows (a_p, k)
    const double* Rcc = M + (int64_t)n * a.ncore * a.nq;             // ncore^2 rows (k, l)
    launch_gemm_nt(Pu, a.nq, nu, a.nq, a.gu, s);
    const int64_t n4 = (int64_t)n * n * n * n;
    hipLaunchKernelGGL(k_df_expand, dim3(grid_of(n4)), dim3(256), 0, s, a.gu, nu, gemm_plan(nu, a.nq).splits, n,
                       a.g_out);
    if (a.ncore > 0)
        hipLaunchKernelGGL(k_df_dvec, dim3(grid_of(a.nq)), dim3(256), 0, s, Rkk, a.ncore, a.nq, a.dvec);
    hipLaunchKernelGGL(k_df_fock, dim3(n * n), dim3(256), 0, s, a.h_full, a.nfull, a.active, n, a.ncore, Pu, a.dvec,
                       M, a.nq, a.h_out);
    if (a.ncore > 0)
        hipLaunchKernelGGL(k_df_rownorm, dim3(a.ncore * a.ncore), dim3(256), 0, s, Rcc, a.nq, a.sq);
    hipLaunchKernelGGL(k_df_efzc, dim3(1), dim3(256), 0, s, a.h_full, a.nfull, a.core, a.ncore, a.dvec, a.sq, a.nq,
                       a.efzc);
}

}  // namespace v2h
