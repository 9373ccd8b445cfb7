// df.hpp -- launcher of the density-fitted integral intake (df.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace v2h {

// Device pointers and sizes of one intake.  The pair list idx (INDEX of each
// row's orbital pair) holds, in order: the n (n + 1) / 2 active pairs
// (a_p, a_q) with p >= q, row p (p + 1) / 2 + q; the ncore pairs (k, k); the
// n ncore pairs (a_p, k), row p ncore + c; the ncore^2 pairs (k, l), row
// c ncore + c'.
// Most partial planes of the pair product (splits of the auxiliary index).
constexpr int kDfMaxSplits = 8;

struct DfArgs {
    const double* qmo;      // [nq][ntri]
    const double* h_full;   // [nfull][nfull], read on and above the diagonal
    const int64_t* idx;     // [m_rows]
    const int* active;      // [n] full-space index of each active orbital
    const int* core;        // [ncore]
    int64_t ntri, m_rows;
    int nfull, nq, n, ncore;
    double* rows;           // [m_rows][nq] scratch
    double* dvec;           // [nq] scratch
    double* gu;             // [kDfMaxSplits][nu][nu] scratch, nu = n (n + 1) / 2
    double* sq;             // [max(ncore^2, 1)] scratch
    double* h_out;          // [n][n]: one-electron part of the source array
    double* g_out;          // [n^2][n^2]: (pq|rs) part of the source array
    double* efzc;           // one double
};

void launch_df_intake(const DfArgs& a, hipStream_t s);

// G[a][b] = sum_Q qmo[Q][a] qmo[Q][b] for a, b < ntri in the 64 x 64 tiles on
// and above the diagonal (so every a <= b); G is ntri x ntri, row-major.
void launch_pair_gram(const double* qmo, int64_t ntri, int nq, double* G, hipStream_t s);

}  // namespace v2h
