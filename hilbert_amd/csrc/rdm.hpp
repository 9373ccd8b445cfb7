// rdm.hpp -- reduced density matrices from a primal vector (host side).
//
// Restates v2RDMSolver::get_opdm / get_tpdm / get_opdm_sparse /
// get_tpdm_sparse (v2rdm_casscf/python_api.cc:39-456) for active spaces with
// no frozen or restricted core (the external-Hamiltonian and Hubbard paths),
// with orbital indices in the user's order.  Unlike the reference, the dense
// 2-RDM is also produced for symmetry-blocked problems.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "layout.hpp"

namespace v2h {

struct OpdmEntry {
    int i, j;
    double value;
};
struct TpdmEntry {
    int i, j, k, l;
    double value;
};

// Dense n x n 1-RDM: type "A", "B" or "SUM" (python_api.cc:39-95, 374-394).
void opdm_dense(const Layout& L, const double* x, const std::string& type, double* out);
// Sparse lists in the reference's order (python_api.cc:39-371).
std::vector<OpdmEntry> opdm_sparse(const Layout& L, const double* x, const std::string& type);
// type "AA", "BB", "AB", "BA" or "SUM"
std::vector<TpdmEntry> tpdm_sparse(const Layout& L, const double* x, const std::string& type);
// Dense spin-free 2-RDM, out[(i n + k) n^2 + j n + l] (python_api.cc:396-456).
void tpdm_dense(const Layout& L, const double* x, double* out);

}  // namespace v2h
