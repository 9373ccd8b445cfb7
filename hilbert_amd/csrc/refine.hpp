// refine.hpp -- warm-started PSD projection of the blocks of 65..256 rows on
// the FP64 matrix cores (refine.hip).
//
// Between two outer iterations of BPSDP, U = mu x + A^T y - c changes little,
// so the eigenvectors of the previous iteration almost diagonalise the new U.
// Each outer iteration refines them with the iteration of Ogita and Aishima
// (Japan J. Indust. Appl. Math. 35, 1007 (2018)), whose steps are matrix
// products only:
//     G = X^T X,  S = X^T U X,  r = I - G,  l_i = s_ii / (1 - r_ii),
//     E_ij = (s_ij + l_j r_ij) / (l_j - l_i)   (|l_j - l_i| above a cluster
//            threshold; r_ij / 2 otherwise),    E_ii = r_ii / 2,
//     X <- X + X E,
// which converges quadratically, restores the orthogonality of X in the same
// step, and then U+ = X diag(l+) X^T.  A block whose correction does not fall
// below the acceptance threshold within max_steps steps, and every block
// without eigenvectors yet (first iteration, after a reset), takes the
// tridiagonal eigensolver instead; its eigenvectors seed the next iteration.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "dbuf.hpp"

namespace v2h {

// One block on the refined path.
struct RefBlock {
    int inst, d, gblock, nt;  // nt: 32-row tiles per side
    int64_t off;              // primal offset of the block inside its instance
    int64_t w;                // work area (doubles): kRefSlots d x d matrices, then d eigenvalues
    int64_t tri;              // the block's tridiagonal work area (reflectors, eigenvectors)
    int64_t part;             // first tile-partial pair (2 doubles per tile)
};
// Matrices of a block's work area, row-major d x d each.
enum RefSlot : int { kSlotV = 0, kSlotU = 1, kSlotXa = 2, kSlotXb = 3, kSlotK = 4, kSlotS = 5, kSlotG = 6 };
constexpr int kRefSlots = 7;

// One 32 x 32 output tile of one block.
struct RefJob {
    int blk, ti, tj, tile;
};

// Per-block state of the refinement.
struct RefState {
    int vvalid;     // slot V holds eigenvector estimates of a recent U
    int phase;      // 0 idle, 1 refining, 2 accepted (at step cstep), 3 rejected
    int cstep;      // step whose correction was accepted
    int xfinal;     // slot of the accepted X
    double max_e;   // largest correction of the last step (separated pairs)
    double cpl;     // largest coupling left inside clusters, relative
    int64_t refined, fallback;  // counts of outer iterations on each path
    // rejections: diverging or non-finite corrections, correction above the
    // acceptance threshold after the last step, clusters left coupled
    int64_t rej_div, rej_e, rej_c;
    // FP64 flops of the block's d x d x d products so far (2 d^3 each: K, G,
    // S and X + XE per step, the final X diag(l+) X^T)
    int64_t flops;
    // After a rejection the block stays on the fallback for `wait` outer
    // iterations (doubling with consecutive rejections, up to kMaxBackoff);
    // its eigenvectors are formed only for the iteration before it tries again.
    int wait, backoff;
    // The instance was running when this iteration's projection began (set
    // by k_ref_state).  The eigenvector formation reads it instead of the
    // instance state, which the next k_pre_reduce on the main stream may
    // change while the formation runs on its side stream.
    int live, pad;
};
constexpr int kMaxBackoff = 64;

struct RefDev {
    const RefBlock* blk;
    const RefJob* jobs;
    RefState* rs;
    int* route;         // [global block]: how this iteration projects the block (kRoute*)
    double* work;
    double* part;
    const double* tri_work;
    int nblk, njobs;
};

// Route of a block of the refined classes in the current outer iteration:
// the tridiagonal fallback known at its start (no eigenvectors, backoff),
// refining, rejected during the refinement (fallback after it), accepted.
enum Route : int { kRouteFallback = 0, kRouteRefining = 1, kRouteRejected = 2, kRouteAccepted = 3 };

// Thresholds of the refinement.
// A step whose largest correction is eps leaves off-diagonal couplings and an
// orthogonality error of order eps^2 (measured: U+ within 7.5e-11 relative of
// LAPACK at eps = 1e-6 on (16e,16o) blocks, so the default keeps it near
// 3e-12).
struct RefParams {
    double cluster_rel = 1e-8;   // |l_i - l_j| <= cluster_rel max|l|: one cluster
    double accept_e = 2e-7;      // accept a step whose largest correction is below this
    double accept_cpl = 1e-13;   // ... and whose clusters are decoupled to this (relative to max|l|)
    // refinement steps before a block is rejected (1..kMaxSteps): 5 keeps
    // most blocks that converge slowly off the tridiagonal fallback, whose
    // launch after the refinement is on the critical path of the iteration
    int max_steps = 5;
};
constexpr int kMaxSteps = 8;

// Start of an outer iteration: phases, backoff and routes of every block.
void launch_ref_state(const BatchDev& B, const RefDev& R, hipStream_t s);
// sym(U) of the refining blocks.
void launch_ref_sym(const BatchDev& B, const RefDev& R, hipStream_t s);
// One refinement step k (1..3): X_in = V (k = 1), Xa (k = 2), Xb (k = 3).
void launch_ref_step(const BatchDev& B, const RefDev& R, const RefParams& P, int k, hipStream_t s);
// x = X diag(l+) X^T / mu with z, v1, partial sums; the block's projection sums.
void launch_ref_final(const BatchDev& B, const RefDev& R, hipStream_t s);
// Eigenvectors of the blocks that took the tridiagonal path this iteration.
void launch_ref_form_v(const BatchDev& B, const RefDev& R, hipStream_t s);
void launch_ref_forget(const RefDev& R, int blk_begin, int blk_end, hipStream_t s);

class Refiner {
  public:
    // items of the tridiagonal classes that take the refined path
    struct Item {
        int inst, d, gblock;
        int64_t off, tri;
    };
    void build(const std::vector<Item>& items, int count, int nblocks_total, hipStream_t s);
    bool empty() const { return blk_.empty(); }
    // One outer iteration: begin() sets every block's route (the fallback
    // launches for kRouteFallback may start right after it); refine() runs
    // the refinement and routes its rejections to kRouteRejected.
    void begin(const BatchDev& B, hipStream_t s) const;
    void refine(const BatchDev& B, hipStream_t s) const;
    // kernel launches of begin() + refine(): state, sym, five per step, final product and close
    int launches() const { return 4 + 5 * params_.max_steps; }
    // After the fallback: eigenvectors of the blocks it projected.
    void seed(const BatchDev& B, hipStream_t s) const;
    void forget(int inst, hipStream_t s) const;
    const int* route() const { return route_.p; }
    // [block projections by the refinement, by the fallback, rejections for
    // divergence, for the correction size, for coupled clusters]
    static constexpr int kStats = 6;
    void stats(int64_t* out) const;
    void set_work_base(const double* tri_work) { tri_work_ = tri_work; }
    void set_params(const RefParams& p) { params_ = p; }
    const RefParams& params() const { return params_; }

  private:
    RefDev dev() const;
    std::vector<RefBlock> blk_;
    std::vector<int> inst_begin_;  // [count + 1] refine blocks of each instance
    int njobs_ = 0;
    RefParams params_;
    DBuf<RefBlock> d_blk_;
    DBuf<RefJob> d_jobs_;
    DBuf<RefState> d_rs_;
    DBuf<int> route_;
    DBuf<double> work_, part_;
    const double* tri_work_ = nullptr;
};

}  // namespace v2h
