// sdp_host.cpp -- BPSDP with caller-supplied maps (see sdp_host.hpp).
#include "sdp_host.hpp"

#include <cmath>
#include <stdexcept>

#include "kernels.hpp"

namespace v2h {

namespace {
double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}
double dot(int64_t n, const double* a, const double* b) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}
}  // namespace

CallbackSdp::CallbackSdp(int64_t n_primal, int64_t n_dual, const SolverOptions& opt, int device)
    : np_(n_primal), nd_(n_dual), opt_(opt), device_(device) {
    if (n_primal <= 0 || n_dual <= 0) throw std::invalid_argument("n_primal and n_dual must be positive");
    int ndev = 0;
    V2H_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw std::invalid_argument("no such GPU");
    V2H_HIP(hipSetDevice(device_));
    V2H_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    y_.assign(nd_, 0.0);
    z_.assign(np_, 0.0);
    u_.assign(np_, 0.0);
    mu_ = opt_.mu0 > 0 ? opt_.mu0 : 1.0;
    d_u_.alloc(np_);
    d_x_.alloc(np_);
    d_z_.alloc(np_);
    d_v1_.alloc(np_);
    d_c_.alloc(np_);
    d_state_.alloc(1);
}

CallbackSdp::~CallbackSdp() {
    if (stream_) (void)hipStreamDestroy(stream_);
}

void CallbackSdp::project(double* x) {
    InstState st{};
    st.mu = mu_;
    st.active = 1;
    const size_t bytes = sizeof(double) * (size_t)np_;
    V2H_HIP(hipMemcpyAsync(d_u_.p, u_.data(), bytes, hipMemcpyHostToDevice, stream_));
    V2H_HIP(hipMemcpyAsync(d_x_.p, x, bytes, hipMemcpyHostToDevice, stream_));
    V2H_HIP(hipMemcpyAsync(d_state_.p, &st, sizeof(st), hipMemcpyHostToDevice, stream_));
    BatchDev B{};
    B.inst = d_inst_.p;
    B.st = d_state_.p;
    B.u = d_u_.p;
    B.x = d_x_.p;
    B.z = d_z_.p;
    B.v1 = d_v1_.p;
    B.c = d_c_.p;
    B.proj_part = proj_.proj_part();
    B.grp = nullptr;
    B.pgroup = -1;  // one instance: every projection launch covers it
    // the warm-started refinement first; the tridiagonal classes take the
    // blocks it did not accept, whose eigenvectors then seed the next call
    const bool refine = proj_.refine_on();
    if (refine) {
        proj_.refine_begin(B, stream_);
        proj_.refine_run(B, stream_);
    }
    for (int k = 0; k < Projector::kClasses; k++) {
        if (!proj_.has(k)) continue;
        if (refine && Projector::refined_class(k)) {
            proj_.launch(k, B, stream_, kRouteFallback);
            proj_.launch(k, B, stream_, kRouteRejected);
        } else {
            proj_.launch(k, B, stream_);
        }
    }
    if (refine) proj_.refine_seed(B, stream_);
    V2H_HIP(hipGetLastError());
    V2H_HIP(hipMemcpyAsync(x, d_x_.p, bytes, hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipMemcpyAsync(z_.data(), d_z_.p, bytes, hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipMemcpyAsync(&st, d_state_.p, sizeof(st), hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
    status |= st.status;
}

int CallbackSdp::solve(double* x, const double* b, const double* c, const int* dims, int nblocks, int maxiter,
                       MapFn Au, MapFn ATu, SdpMonitorFn mon, int print_level, void* data) {
    if (!x || !b || !c || !dims || nblocks <= 0 || !Au || !ATu) throw std::invalid_argument("null argument");
    V2H_HIP(hipSetDevice(device_));
    int64_t tot = 0;
    for (int k = 0; k < nblocks; k++) {
        if (dims[k] < 0) throw std::invalid_argument("negative block size");
        tot += (int64_t)dims[k] * dims[k];
    }
    if (tot != np_) throw std::invalid_argument("block sizes do not add up to n_primal");
    std::vector<int> d(dims, dims + nblocks);
    if (d != dims_) {
        dims_ = d;
        inst_.assign(1, InstDev{});
        proj_.build({dims_}, inst_, stream_);
        d_inst_.upload(inst_);
    }
    if (!started_) {
        std::fill(y_.begin(), y_.end(), 0.0);
        std::fill(z_.begin(), z_.end(), 0.0);
        mu_ = opt_.mu0 > 0 ? opt_.mu0 : 1.0;
        eps_p = eps_d = 1.0;
        started_ = true;
    }
    V2H_HIP(hipMemcpyAsync(d_c_.p, c, sizeof(double) * (size_t)np_, hipMemcpyHostToDevice, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));

    double bn = std::sqrt(dot(nd_, b, b)), cn = std::sqrt(dot(np_, c, c));
    if (bn == 0.0) bn = 1.0;
    if (cn == 0.0) cn = 1.0;
    std::vector<double> Ax(nd_), rhs(nd_), r(nd_), p(nd_), w(nd_), tp(np_), aty(np_);
    converged = 0;
    Au(Ax.data(), x, data);
    for (int it = 0; it < maxiter; it++) {
        if (opt_.maxiter > 0 && oiter >= opt_.maxiter) break;
        // rhs = A (c - z) + mu (b - A x); r = rhs - A A^T y
        for (int64_t i = 0; i < np_; i++) tp[i] = c[i] - z_[i];
        Au(rhs.data(), tp.data(), data);
        for (int64_t i = 0; i < nd_; i++) rhs[i] += mu_ * (b[i] - Ax[i]);
        ATu(tp.data(), y_.data(), data);
        Au(w.data(), tp.data(), data);
        for (int64_t i = 0; i < nd_; i++) r[i] = p[i] = rhs[i] - w[i];
        double rr = dot(r, r);
        double e = oiter == 0 ? 1.0 : std::max(eps_p, eps_d);
        double tol = std::min(kCgTolFactor * e, kCgTolCap);
        if (tol < opt_.cg_convergence) tol = opt_.cg_convergence;
        const double thr = tol * tol * dot(rhs, rhs);
        int k = 0;
        while (k < opt_.cg_maxiter && rr > thr) {
            ATu(tp.data(), p.data(), data);
            Au(w.data(), tp.data(), data);
            const double pw = dot(p, w);
            if (pw <= 0.0) break;
            const double al = rr / pw;
            for (int64_t i = 0; i < nd_; i++) {
                y_[i] += al * p[i];
                r[i] -= al * w[i];
            }
            const double rr1 = dot(r, r);
            const double be = rr1 / rr;
            rr = rr1;
            for (int64_t i = 0; i < nd_; i++) p[i] = r[i] + be * p[i];
            k++;
        }
        iiter += k;
        last_cg = k;
        // U = mu x + A^T y - c, projected block by block on the GPU
        ATu(aty.data(), y_.data(), data);
        for (int64_t i = 0; i < np_; i++) u_[i] = mu_ * x[i] + aty[i] - c[i];
        project(x);
        if (status) break;
        double ed = 0.0;
        for (int64_t i = 0; i < np_; i++) {
            const double t = aty[i] - c[i] + z_[i];
            ed += t * t;
        }
        Au(Ax.data(), x, data);
        double ep = 0.0;
        for (int64_t i = 0; i < nd_; i++) {
            const double t = Ax[i] - b[i];
            ep += t * t;
        }
        eps_p = std::sqrt(ep) / bn;
        eps_d = std::sqrt(ed) / cn;
        energy_primal = dot(np_, c, x);
        energy_dual = dot(nd_, b, y_.data());
        oiter++;
        if (opt_.mu_update_frequency > 0 && oiter % opt_.mu_update_frequency == 0 && eps_d > 0.0 && eps_p > 0.0)
            mu_ *= eps_p / eps_d;
        if (mon) mon(print_level, (int)oiter, k, energy_primal, energy_dual, mu_, eps_p, eps_d, data);
        if (std::fabs(energy_primal - energy_dual) < opt_.e_convergence && eps_p < opt_.r_convergence &&
            eps_d < opt_.r_convergence) {
            converged = 1;
            break;
        }
    }
    return converged;
}

}  // namespace v2h
