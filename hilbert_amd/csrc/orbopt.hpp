// orbopt.hpp -- device kernels of the v2RDM-CASSCF orbital step (orbopt.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "layout.hpp"

namespace v2h {

// C[b] = alpha op(A[b]) op(B[b]) + beta C[b], row-major, b < batch; op(A) is
// M x K (A stored K x M with lda when ta), op(B) is K x N (B stored N x K
// when tb).  Strides sA / sB / sC between batch entries (0: shared).
// dg is added on the diagonal of every result (m == n).
void launch_dgemm(bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int64_t lda, int64_t sA,
                  const double* B, int64_t ldb, int64_t sB, double beta, double* C, int64_t ldc, int64_t sC,
                  int batch, hipStream_t s, double dg = 0.0);
// One product with K split over up to kMaxSlices workgroup slices (partial
// products in work, [slices][M][N], then summed in slice order): for a small
// M x N with a long K.
constexpr int kMaxSlices = 64;
// C = alpha op(A) B (+ dg on the diagonal), beta = 0, one product; the LDS
// kernel when M, N, K <= 128, else launch_dgemm.
void launch_dgemm_small(bool ta, int M, int N, int K, double alpha, const double* A, int64_t lda, const double* B,
                        int64_t ldb, double* C, int64_t ldc, hipStream_t s, double dg = 0.0);
int dgemm_split_slices(int M, int N, int K);
void launch_dgemm_split(bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* work,
                        hipStream_t s);

// dense [nq][nfull][nfull] <-> packed lower triangle [nq][nfull (nfull + 1) / 2]
void launch_df_unpack(const double* qmo, int nq, int nfull, double* bd, hipStream_t s);
void launch_df_pack(const double* bd, int nq, int nfull, double* qmo, hipStream_t s);

// Layout lookups of the active-space densities, Pitzer order.
struct OoTables {
    int n, ncore, no;
    const int4* pair;        // [n][n]: (irrep of the pair, ab index, aa index or -1, +-1 order sign)
    const int64_t* d1;       // [2][n][n]: offsets of D1a(i, j) and D1b(i, j), -1 across irreps
    int64_t offab[kMaxIrrep], offaa[kMaxIrrep], offbb[kMaxIrrep];
    int gab[kMaxIrrep], gaa[kMaxIrrep];
};

// D [no][no] and the symmetrised spin-free 2-RDM d [no^4] over O = core +
// active (Pitzer) from one instance's primal vector x.
void launch_oo_density(const OoTables& t, const double* x, double* D, double* d, hipStream_t s);

// BO[Q][r][s] = B[Q][occ r][occ s], BFO[p][r][Q] = B[Q][p][occ r], Bdiag[p][Q] = B[Q][p][p]
void launch_oo_gather(const double* bd, int nq, int nfull, const int* occ, int no, double* BO, double* BFO,
                      double* Bdiag, hipStream_t s);

// F[x][y] += sum_r D[x][r] h[y][occ r]
void launch_oo_fock1(const double* D, const double* h, const int* occ, int no, int nfull, double* F,
                     hipStream_t s);

struct OoGradArgs {
    const int2* pairs;       // [nrot] (p, q) full-space indices
    int nrot, nfull, no;
    const int* pos;          // [nfull]: O position or -1
    const double *F, *D, *d, *h, *gO, *Jd, *Kd;
    double *g, *H;           // [nrot]
};
void launch_oo_grad_hess(const OoGradArgs& a, hipStream_t s);

// Precondition_Step + scaling: kref = -g / max(|H|, 1e-2); kcur = ksave =
// r kref.  Before that, out[0] = g.g and out[1] = g.kcur + 1/2 sum kcur H kcur
// with the previous kcur (one workgroup).
void launch_oo_precondition(const double* g, const double* H, int nrot, double r, double* kref, double* kcur,
                            double* ksave, double* out, hipStream_t s);
// Update_kappa: kcur = r kref - kcur; ksave += kcur
void launch_oo_update_kappa(const double* kref, int nrot, double r, double* kcur, double* ksave, hipStream_t s);

// K = 0; K[p][q] = scale kappa, K[q][p] = -scale kappa; out[0] = max column
// abs sum of K (one workgroup for the norm).
void launch_oo_kappa_matrix(const double* kappa, const int2* pairs, int nrot, double scale, int nfull, double* K,
                            double* out, hipStream_t s);
void launch_oo_identity(double* U, int nfull, hipStream_t s);
// y += a x
void launch_oo_axpy(double a, const double* x, double* y, int64_t n, hipStream_t s);
// y = x - z
void launch_oo_sub(const double* x, const double* z, double* y, int64_t n, hipStream_t s);

}  // namespace v2h
