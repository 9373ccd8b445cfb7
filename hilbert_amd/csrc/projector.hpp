// projector.hpp -- the PSD projection of every block of a batch: blocks
// grouped by size class, one launch per class and outer iteration, and the
// device work areas the classes need.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "dbuf.hpp"

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
    void launch(int k, const BatchDev& B, hipStream_t s) const;
    int64_t nblocks() const { return nblocks_; }
    double* proj_part() const { return proj_part_.p; }

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
};

}  // namespace v2h
