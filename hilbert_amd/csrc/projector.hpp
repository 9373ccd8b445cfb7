// projector.hpp -- the PSD projection of every block of a batch: blocks
// grouped by size class, one launch per class and outer iteration, and the
// device work areas the classes need.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "dbuf.hpp"
#include "refine.hpp"

namespace v2h {

class Projector {
  public:
    // Jacobi d <= 16, tridiagonal d <= 64 / 128 / 192 / 224 / 256, Jacobi d > 256
    static constexpr int kClasses = 7;

    // dims[i]: PSD block sizes of instance i in primal order (zeros allowed);
    // fills block_first / nblocks of inst[i].  Blocks are numbered
    // consecutively (the index of a block's partial sums in proj_part()).
    void build(const std::vector<std::vector<int>>& dims, std::vector<InstDev>& inst, hipStream_t s);
    // forget instance i's eigenvalues (the warm start of the tridiagonal path)
    void forget(int i, hipStream_t s);
    bool has(int k) const { return !cls_[k].items.empty(); }
    // The launch of class k.  With the refinement on, a refined class (2..5)
    // is launched twice per outer iteration: route kRouteFallback (blocks known
    // to need the tridiagonal eigensolver when the iteration starts, alongside
    // the refinement) and kRouteRejected (after it); route < 0: every block.
    void launch(int k, const BatchDev& B, hipStream_t s, int route = -1) const;
    int64_t nblocks() const { return nblocks_; }
    double* proj_part() const { return proj_part_.p; }

    // Warm-started refinement (refine.hpp) for the tridiagonal classes of
    // 65..256 rows: on by default.
    static bool refined_class(int k) { return k >= 2 && k <= 5; }
    bool refine_on() const { return refine_on_ && !refiner_.empty(); }
    void set_refine(bool on) { refine_on_ = on; }
    void refine_begin(const BatchDev& B, hipStream_t s) const { refiner_.begin(B, s); }
    void refine_run(const BatchDev& B, hipStream_t s) const { refiner_.refine(B, s); }
    void refine_seed(const BatchDev& B, hipStream_t s) const { refiner_.seed(B, s); }
    void refine_stats(int64_t* out) const { refiner_.stats(out); }
    void set_refine_params(const RefParams& p) { refiner_.set_params(p); }
    const RefParams& refine_params() const { return refiner_.params(); }
    int refine_launches() const { return refiner_.launches(); }

  private:
    struct Class {
        int nt = 0, max_d = 0;
        bool global = false;
        int tri = -1;  // tridiagonal kernel class, or -1 for Jacobi
        std::vector<ProjItem> items;
        DBuf<ProjItem> d_items;
    };
    Class cls_[kClasses];
    DBuf<double> proj_part_, rec_, gmat_, work_, lam_;
    std::vector<int64_t> lam_begin_;  // [count + 1] eigenvalue-store range of each instance
    int64_t nblocks_ = 0;
    int max_sweeps_ = 24;
    Refiner refiner_;
    bool refine_on_ = true;
};

}  // namespace v2h
