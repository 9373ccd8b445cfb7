// sdp_host.hpp -- BPSDP with caller-supplied constraint maps (libsdp's
// generic contract: SDPSolver::solve(x, b, c, block_dims, maxiter, Au, ATu,
// monitor, ..., data), v2rdm_solver.cc:1340 and v2rdm_doci_solver.cc).
//
// The maps are host functions, so the CG vectors live in host memory next to
// them; the PSD projection of the blocks runs on the GPU (the same kernels as
// the DQG batch path).  Iteration, stopping rule and mu update are those of
// the batch solver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "dbuf.hpp"
#include "projector.hpp"

namespace v2h {

using MapFn = void (*)(double* out, double* in, void* data);
// libsdp's SDPProgressMonitorFunction (v2rdm_solver.cc:115-124): print_level
// first, as sdp_->solve passes it through (v2rdm_solver.cc:1340).
using SdpMonitorFn = void (*)(int print_level, int oiter, int iiter, double energy_primal, double energy_dual,
                              double mu, double primal_error, double dual_error, void* data);

class CallbackSdp {
  public:
    CallbackSdp(int64_t n_primal, int64_t n_dual, const SolverOptions& opt, int device);
    ~CallbackSdp();
    CallbackSdp(const CallbackSdp&) = delete;
    CallbackSdp& operator=(const CallbackSdp&) = delete;

    void set_options(const SolverOptions& o) { opt_ = o; }
    // Up to maxiter outer iterations from x (overwritten with the result);
    // y, z and mu persist across calls.  mon, if set, is called after every
    // outer iteration with print_level passed through.  Returns 1 when
    // converged.
    int solve(double* x, const double* b, const double* c, const int* dims, int nblocks, int maxiter, MapFn Au,
              MapFn ATu, SdpMonitorFn mon, int print_level, void* data);

    int64_t n_primal() const { return np_; }
    int64_t n_dual() const { return nd_; }
    double mu() const { return mu_; }
    void set_mu(double m) { mu_ = m; }
    const std::vector<double>& y() const { return y_; }
    const std::vector<double>& z() const { return z_; }
    void set_y(const double* y) { y_.assign(y, y + nd_); started_ = true; }
    void set_z(const double* z) { z_.assign(z, z + np_); started_ = true; }
    double energy_primal = 0.0, energy_dual = 0.0, eps_p = 1.0, eps_d = 1.0;
    int64_t oiter = 0, iiter = 0;
    int converged = 0, status = 0, last_cg = 0;

  private:
    void project(double* x);  // x <- U+ / mu, z <- U+ - sym(U) for u_ on the device

    int64_t np_, nd_;
    SolverOptions opt_;
    int device_;
    hipStream_t stream_ = nullptr;
    bool started_ = false;
    double mu_ = 1.0;
    std::vector<double> y_, z_, u_;
    std::vector<int> dims_;
    Projector proj_;
    std::vector<InstDev> inst_;
    DBuf<InstDev> d_inst_;
    DBuf<InstState> d_state_;
    DBuf<double> d_u_, d_x_, d_z_, d_v1_, d_c_;
};

}  // namespace v2h
