// df.hpp -- launcher of the density-fitted integral intake (df.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace v2h {

// Device pointers and sizes of one intake.  The pair list idx (INDEX of each
// row's orbital pair) holds, in order: the n^2 active pairs (a_p, a_q), row
// p n + q; the ncore pairs (k, k); the n ncore pairs (a_p, k), row p ncore + c;
// the ncore^2 pairs (k, l), row c ncore + c'.
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
    double* kmat;           // [n][n] scratch
    double* h_out;          // [n][n]: one-electron part of the source array
    double* g_out;          // [n^2][n^2]: (pq|rs) part of the source array
    double* efzc;           // one double
};

void launch_df_intake(const DfArgs& a, hipStream_t s);

}  // namespace v2h
