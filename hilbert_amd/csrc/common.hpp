// common.hpp -- device-resident data structures shared by the host driver and
// the HIP kernels.  One "structure" holds the gather tables of one problem
// shape (shared by every instance with the same shape); one "instance" is one
// SDP of the batch, whose vectors live at fixed offsets inside batch-wide
// primal and dual arrays.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#define V2H_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw v2h::HipError(std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

namespace v2h {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

constexpr int kLongFlag = 1 << 30;   // work-unit code: long row (else: slice)

// SELL-64 operator tables on the device (see sell.hpp).
struct SellDev {
    const uint32_t* ent;
    const uint32_t* slice_start;
    const uint32_t* long_start;
    const uint32_t* long_ent;
    const int32_t* long_row;
    const double* coef;     // [256]
    int64_t nrows;
    int nslices, nlong;
};

struct StructDev {
    SellDev A, AT;
    const double* b;        // [n_dual]
    const int32_t* cost1;   // [n_primal] cost-vector gather map
    const int32_t* cost2;
    const int32_t* adj_ptr; // [nsrc + 1] transposed cost map (adjoint gather)
    const int32_t* adj_ent; // primal index, sign in bit 31 (set = minus)
    int64_t n_primal, n_dual, nsrc;
};

struct InstDev {
    int64_t poff, doff;     // offsets into batch primal / dual arrays
    int64_t soff;           // offset into the batch integral array [h; g]
    int sid;                // structure id
    int block_first, nblocks;
};

// Per-instance solver state, written only by the reducing workgroup of a
// kernel (the last one to finish) or by the host between launches.
struct InstState {
    double mu, bnorm, cnorm;
    double e_primal, e_dual, eps_p, eps_d;
    double cg_thr;              // tau^2 ||rhs||^2 of the current CG solve
    double rr, pw, rw, ww;      // CG scalars of the last matvec
    int64_t oiter, iiter;
    int cg_iter, cg_done, cg_last;
    int active, converged, pending, status;  // pending: a projection awaits its bookkeeping
    int maxiter_left;           // outer iterations left in the current solve() call
};

enum StatusBits : int { kStatusJacobiNotConverged = 1, kStatusNonFinite = 2 };

// CG stopping rule of each outer iteration: relative residual below
// kCgTolFactor x the larger of the last primal and dual errors, capped at
// 1e-2 first and then raised to at least CG_CONVERGENCE.  Shared by the batch
// kernels and the caller-map solver.
#ifndef V2H_CG_TOL_FACTOR
#define V2H_CG_TOL_FACTOR 0.003
#endif
constexpr double kCgTolFactor = V2H_CG_TOL_FACTOR;
constexpr double kCgTolCap = 1e-2;

struct SolverOptions {
    double e_convergence, r_convergence, cg_convergence, mu0;
    int cg_maxiter, maxiter, mu_update_frequency;
    double cg_tol_factor = kCgTolFactor;  // the factor of the CG stopping rule above
};

// CG state of one dual row, stored together so that a gather of the row
// touches one 32-byte sector: residual r, w = A A^T p, direction p, dual y.
struct alignas(32) CgRow {
    double r, w, p, y;
};

// Batch-wide device arrays.
struct BatchDev {
    const StructDev* structs;
    const InstDev* inst;
    InstState* st;
    // primal-length arrays
    double *x, *z, *c, *v1, *u, *aty, *q;
    // dual-length arrays
    CgRow* cg;
    double *av1, *ax, *aaty;
    // reductions
    double* partials;       // per-workgroup partial dot products
    double* proj_part;      // per-block projection partials (c.x, dual residual)
    SolverOptions opt;
    // Instances a projection launch covers: every active one (pgroup < 0, the
    // default of a value-initialised BatchDev), or the active ones with
    // grp[inst] == pgroup (the pipelined solve projects one group of
    // instances while the other group runs its CG).
    const int* grp = nullptr;
    int pgroup = -1;
};
static_assert(BatchDev{}.pgroup < 0 && BatchDev{}.grp == nullptr, "a value-initialised BatchDev covers every instance");

__device__ __forceinline__ bool in_pgroup(const BatchDev& B, int inst) {
    return B.pgroup < 0 || (B.grp != nullptr && B.grp[inst] == B.pgroup);
}

// One wavefront's share of a gather launch over one instance, with what the
// wavefront needs to start loading.  Slices: width x 64 entries from ent,
// k-major; meta = width | (valid lanes << 16); row = first row.  Long rows:
// meta = n | kRecLong, n entries from ent, row = the row.  meta = 0: no work.
constexpr uint32_t kRecLong = 1u << 31;
struct alignas(16) UnitRec {
    const uint32_t* ent;
    const double* coef;     // [256] coefficient table of the operator
    uint32_t meta;
    int32_t row;
};

// One work unit list per operator direction.
struct UnitsDev {
    const UnitRec* rec;     // 4 per workgroup
    const int* wg_inst;     // instance of each workgroup
    const int* wg_begin;    // [ninst + 1] first workgroup of each instance
    int nwg;
    int ninst;
    // Launch plan: the launched workgroups are runs of logical workgroups
    // (the workgroups of the instances still iterating).  plan[0..plan_m] are
    // the prefix counts of the runs, plan[plan_m + 1 + j] the first logical
    // workgroup of run j.  plan == nullptr launches every workgroup.
    const int* plan;
    int plan_m;
    int grid;               // workgroups launched
    int remap;              // 1: blocks dealt to the XCDs in contiguous ranges of logical workgroups
};

struct ProjItem {
    int inst, d, gblock, pad;
    int64_t off;            // primal offset inside the instance
    int64_t rec_off;        // offset of the rotation record (doubles)
    int64_t mat_off;        // offset of the packed matrix when it lives in global memory
    int64_t work_off;       // tridiagonal path: offset of the block's work area (7 d^2 doubles)
    int64_t lam_off;        // tridiagonal path: offset of the block's eigenvalues (d + 1 doubles)
};

}  // namespace v2h
