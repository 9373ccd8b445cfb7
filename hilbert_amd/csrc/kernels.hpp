// kernels.hpp -- host-side launchers of the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace v2h {

void launch_cg_at(const BatchDev& B, const UnitsDev& U, hipStream_t s);
void launch_cg_a(const BatchDev& B, const UnitsDev& U, hipStream_t s);
void launch_pre(const BatchDev& B, const UnitsDev& U, hipStream_t s);
void launch_post(const BatchDev& B, const UnitsDev& U, hipStream_t s);
void launch_spmv_plain(const SellDev& S, const int4* units, int wg0, int nwg, const double* in, double* out,
                       hipStream_t s);

// LDS bytes of the Jacobi projection for block size d and nt threads (the
// packed matrix is in LDS unless global_matrix).
size_t project_lds_bytes(int d, int nt, bool global_matrix);
void launch_project(int nt, bool global_matrix, const BatchDev& B, const ProjItem* items, int nitems, int max_d,
                    double* rec, int max_sweeps, double* gmat, hipStream_t s);
void set_project_lds_limit(size_t bytes);

// Tridiagonal projection (eig_tri.hip), classes 0..4 for blocks up to
// 64 / 128 / 192 / 224 / 256.  work: 7 d^2 doubles per block (class 4: plus
// 1024 x 64 for the tile); lam: d + 1 doubles per block, kept between calls
// (NaN = no previous eigenvalues).  route (may be null, classes 1-4): only
// blocks with route[global block] == want are projected.  Each reflector's tau
// is kept on the diagonal of the reflector matrix (row k, column k).
int project_tri_max_d(int cls);
void launch_project_tri(int cls, const BatchDev& B, const ProjItem* items, int nitems, double* work, double* lam,
                        const int* route, int want, hipStream_t s);

void launch_fill(double* p, int64_t n, double v, hipStream_t s);
void launch_copy(double* dst, const double* src, int64_t n, hipStream_t s);
// dst[i * dst_stride] = src[i * src_stride], i < n (strides in doubles)
void launch_copy_strided(double* dst, int64_t dst_stride, const double* src, int64_t src_stride, int64_t n,
                         hipStream_t s);
void launch_pack_cost(const int32_t* i1, const int32_t* i2, const double* src, double* c, int64_t n,
                      hipStream_t s);
void launch_cost_adjoint(const int32_t* ptr, const int32_t* ent, const double* xbar, double* dsrc, int64_t nsrc,
                         hipStream_t s);
// out[0] = a . b; out[1..256] receive the partial sums
void launch_dot1(const double* a, const double* b, int64_t n, double* out, hipStream_t s);

}  // namespace v2h
