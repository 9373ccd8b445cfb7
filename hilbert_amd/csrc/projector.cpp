// projector.cpp -- grouping of PSD blocks by size class and their launches.
#include "projector.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "kernels.hpp"

namespace v2h {

namespace {
// Projection classes by block size.  Blocks of 17..256 take the tridiagonal
// path (one wavefront per block up to 64, register tiles up to 192,
// registers + LDS up to 224, the tile in global memory up to 256); the smallest blocks and those above 256 take
// parallel Jacobi (packed matrix in LDS for d <= 16, in global memory beyond
// 256).
constexpr int kGlobalClass = 6;
constexpr size_t kLdsLimit = 160 * 1024;
constexpr int kClassThreads[Projector::kClasses] = {64, 64, 512, 768, 1024, 1024, 1024};
constexpr int kClassTri[Projector::kClasses] = {-1, 0, 1, 2, 3, 4, -1};
int proj_class(int d) {
    if (d <= 16) return 0;
    if (d <= 64) return 1;
    if (d <= 128) return 2;
    if (d <= 192) return 3;
    if (d <= 224) return 4;
    if (d <= 256) return 5;
    return kGlobalClass;
}
// work doubles of a tridiagonal-path block: reflectors, eigenvectors, LU
// factors (7 d^2), plus the tile of the global-memory class
int64_t tri_work(int tri, int d) { return 7 * (int64_t)d * d + (tri == 4 ? (int64_t)1024 * 8 * 8 : 0); }
}  // namespace

void Projector::build(const std::vector<std::vector<int>>& dims, std::vector<InstDev>& inst, hipStream_t s) {
    int gblock = 0;
    int64_t rec = 0, mat = 0, work = 0, lam = 0;
    const int count = (int)dims.size();
    std::vector<Refiner::Item> ritems;
    for (auto& c : cls_) c.items.clear();
    lam_begin_.assign(count + 1, 0);
    for (int i = 0; i < count; i++) {
        inst[i].block_first = gblock;
        lam_begin_[i] = lam;
        int nb = 0;
        int64_t off = 0;
        for (const int dim : dims[i]) {
            const int64_t boff = off;
            off += (int64_t)dim * dim;
            if (dim == 0) continue;
            const int k = proj_class(dim);
            // the Jacobi kernels give each rotation pair one thread
            if (kClassTri[k] < 0 && (dim + 1) / 2 > kClassThreads[k])
                throw std::invalid_argument("PSD block larger than 2048 is not supported");
            ProjItem it{};
            it.inst = i;
            it.d = dim;
            it.gblock = gblock++;
            it.off = boff;
            if (kClassTri[k] >= 0) {
                it.work_off = work;
                work += tri_work(kClassTri[k], dim);
                it.lam_off = lam;
                lam += dim + 1;
                if (refined_class(k)) ritems.push_back(Refiner::Item{i, dim, it.gblock, boff, it.work_off});
            } else {
                int dp = dim + (dim & 1);
                it.rec_off = rec;
                rec += (int64_t)max_sweeps_ * (dp - 1) * (dp / 2) * 2;
                if (k == kGlobalClass) {
                    it.mat_off = mat;
                    mat += (int64_t)dp * (dp + 1) / 2;
                }
            }
            cls_[k].items.push_back(it);
            nb++;
        }
        inst[i].nblocks = nb;
    }
    lam_begin_[count] = lam;
    nblocks_ = gblock;
    size_t lds_max = 0;
    for (int k = 0; k < kClasses; k++) {
        auto& c = cls_[k];
        c.nt = kClassThreads[k];
        c.global = k == kGlobalClass;
        c.tri = kClassTri[k];
        std::stable_sort(c.items.begin(), c.items.end(), [](const ProjItem& a, const ProjItem& b) { return a.d > b.d; });
        c.max_d = c.items.empty() ? 0 : c.items.front().d;
        if (!c.items.empty() && c.tri < 0) {
            size_t lds = project_lds_bytes(c.max_d, c.nt, c.global);
            if (lds > kLdsLimit) throw std::invalid_argument("PSD block does not fit in LDS");
            lds_max = std::max(lds_max, lds);
        }
        if (!c.items.empty() && c.tri >= 0 && c.max_d > project_tri_max_d(c.tri))
            throw std::logic_error("projection class table is inconsistent");
        c.d_items.upload(c.items);
    }
    if (lds_max > 64 * 1024) set_project_lds_limit(lds_max);
    proj_part_.alloc(2 * (size_t)std::max(gblock, 1));
    rec_.alloc(std::max<int64_t>(rec, 1));
    gmat_.alloc(std::max<int64_t>(mat, 1));
    work_.alloc(std::max<int64_t>(work, 1));
    lam_.alloc(std::max<int64_t>(lam, 1));
    launch_fill(lam_.p, std::max<int64_t>(lam, 1), std::nan(""), s);
    refiner_.build(ritems, count, gblock, s);
    refiner_.set_work_base(work_.p);
}

void Projector::forget(int i, hipStream_t s) {
    if (lam_begin_[i + 1] > lam_begin_[i])
        launch_fill(lam_.p + lam_begin_[i], lam_begin_[i + 1] - lam_begin_[i], std::nan(""), s);
    refiner_.forget(i, s);
}

void Projector::launch(int k, const BatchDev& B, hipStream_t s, int route) const {
    const Class& c = cls_[k];
    if (c.items.empty()) return;
    if (c.tri >= 0)
        launch_project_tri(c.tri, B, c.d_items.p, (int)c.items.size(), work_.p, lam_.p,
                           route >= 0 && refined_class(k) && refine_on() ? refiner_.route() : nullptr, route, s);
    else
        launch_project(c.nt, c.global, B, c.d_items.p, (int)c.items.size(), c.max_d, rec_.p, max_sweeps_, gmat_.p, s);
}

}  // namespace v2h
