// constraints.hpp -- the DQG + spin constraint matrix A (rows in the order of
// hilbert's bpsdp_Au) and the right-hand side b, built on the host once per
// problem structure.  The device only ever sees the packed gather tables.
#pragma once

#include <cstdint>
#include <vector>

#include "layout.hpp"

namespace v2h {

// Compressed sparse rows with one right-hand-side value per row.
struct Csr {
    int64_t nrows = 0, ncols = 0;
    std::vector<int64_t> ptr;  // [nrows + 1]
    std::vector<int32_t> col;
    std::vector<double> val;
    std::vector<double> rhs;   // b (empty for a transpose)
    int64_t nnz() const { return (int64_t)col.size(); }
    Csr transpose() const;
};

// Rows: traces, hermiticity, D1/Q1, contractions (d2.cc:46-594); spin
// (spin.cc:46-508); Q2 (q2.cc:170-438); G2 (g2.cc:245-634); b from
// v2rdm_solver.cc:1826-2074.  Duplicate columns in a row are summed and
// exact zeros dropped.
Csr build_constraints(const Layout& L);

// TPDM_GUESS HF (v2rdm_solver.cc:1711-1798 with Q2_constraints_guess,
// q2.cc:45-167, and G2_constraints_guess, g2.cc:44-242): D1, Q1, D2ab/aa/bb
// of the determinant that occupies the first nocc_alpha / nocc_beta orbitals
// of the caller's order, and every Q2 / G2 element from the constraint row
// that defines it (the rows are linear in D1, Q1, D2, so this is the same
// value the reference's guess loops compute).  D200 stays zero, as in the
// reference.
std::vector<double> hf_guess(const Layout& L, const Csr& A, int nocc_alpha, int nocc_beta);

// Gather map of the cost vector: c[e] = src[i1[e]] - src[i2[e]] with
// src = [h (n*n, user order) ; g (n^4, user order)] and index -1 meaning 0
// (v2rdm_solver.cc:164-193, tei.cc:139-199).
struct CostMap {
    std::vector<int32_t> i1, i2;
};
CostMap build_cost_map(const Layout& L);

}  // namespace v2h
