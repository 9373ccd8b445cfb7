// abi.cpp -- extern "C" entry points (include/v2rdm_hip.h).  Exceptions are
// caught here and turned into status codes plus a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/v2rdm_hip.h"
#include "dbuf.hpp"
#include "kernels.hpp"
#include "noalloc.hpp"
#include "rdm.hpp"
#include "sdp_host.hpp"
#include "solver.hpp"
#include "writers.hpp"

struct v2h_batch {
    std::unique_ptr<v2h::Batch> impl;
};

namespace {

thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
    try {
        f();
        return V2H_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return V2H_ERR_INVALID_ARGUMENT;
    } catch (const v2h::HipError& e) {
        g_err = e.what();
        return V2H_ERR_HIP;
    } catch (const std::bad_alloc& e) {
        g_err = "out of host memory";
        return V2H_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return V2H_ERR_INTERNAL;
    } catch (...) {
        g_err = "unknown error";
        return V2H_ERR_INTERNAL;
    }
}

v2h::SolverOptions to_opts(const v2h_sdp_options* o) {
    v2h::SolverOptions s{};
    s.e_convergence = o->sdp_objective_convergence;
    s.r_convergence = o->sdp_error_convergence;
    s.cg_convergence = o->cg_convergence;
    s.cg_maxiter = o->cg_maxiter;
    s.maxiter = o->maxiter;
    s.mu_update_frequency = o->mu_update_frequency;
    s.mu0 = o->mu0;
    return s;
}

v2h::Batch& B(v2h_batch* b) {
    if (!b || !b->impl) throw std::invalid_argument("null batch handle");
    return *b->impl;
}
const v2h::Batch& B(const v2h_batch* b) {
    if (!b || !b->impl) throw std::invalid_argument("null batch handle");
    return *b->impl;
}
void check_inst(const v2h::Batch& b, int i) {
    if (i < 0 || i >= b.count()) throw std::invalid_argument("instance index out of range");
}
void check_ptr(const void* p) {
    if (!p) throw std::invalid_argument("null pointer argument");
}

}  // namespace

extern "C" {

const char* v2h_last_error(void) { return g_err.c_str(); }

int v2h_default_options(v2h_sdp_options* o) {
    return guarded([&] {
        check_ptr(o);
        o->sdp_objective_convergence = 1e-6;
        o->sdp_error_convergence = 1e-5;
        o->cg_convergence = 1e-9;
        o->cg_maxiter = 10000;
        o->maxiter = 50;
        o->mu_update_frequency = 1000;
        o->mu0 = 1.0;
    });
}

int v2h_device_count(int* count) {
    return guarded([&] {
        check_ptr(count);
        V2H_HIP(hipGetDeviceCount(count));
    });
}

int v2h_calibration_copy(int device, int64_t n, int repeats, double* ms_per_copy) {
    return guarded([&] {
        check_ptr(ms_per_copy);
        if (n <= 0 || repeats <= 0) throw std::invalid_argument("n and repeats must be positive");
        V2H_HIP(hipSetDevice(device));
        v2h::DBuf<double> src, dst;
        src.alloc(n);
        dst.alloc(n);
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        V2H_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        V2H_HIP(hipEventCreate(&e0));
        V2H_HIP(hipEventCreate(&e1));
        v2h::launch_fill(src.p, n, 1.0, s);
        v2h::launch_fill(dst.p, n, 0.0, s);
        V2H_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < repeats; r++) v2h::launch_copy(dst.p, src.p, n, s);
        V2H_HIP(hipEventRecord(e1, s));
        V2H_HIP(hipGetLastError());
        V2H_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        V2H_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_copy = ms / repeats;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    });
}

int v2h_batch_create(const v2h_problem_desc* descs, int count, const v2h_sdp_options* opt, int device,
                     v2h_batch** out) {
    return guarded([&] {
        check_ptr(descs);
        check_ptr(opt);
        check_ptr(out);
        *out = nullptr;
        if (count <= 0) throw std::invalid_argument("count must be positive");
        std::vector<v2h::ProblemDesc> ds(count);
        for (int i = 0; i < count; i++) {
            const v2h_problem_desc& s = descs[i];
            v2h::ProblemDesc& d = ds[i];
            d.nmo = s.nmo;
            d.nalpha = s.nalpha;
            d.nbeta = s.nbeta;
            d.multiplicity = s.multiplicity;
            d.nirrep = s.nirrep;
            if (s.orbital_irreps) d.irreps.assign(s.orbital_irreps, s.orbital_irreps + s.nmo);
            d.q2 = s.constrain_q2 != 0;
            d.g2 = s.constrain_g2 != 0;
            d.spin = s.constrain_spin != 0;
            d.sz = s.constrain_sz != 0;
            d.t1 = s.constrain_t1 != 0;
            d.t2 = s.constrain_t2 != 0;
            d.d3 = s.constrain_d3 != 0;
        }
        auto h = std::make_unique<v2h_batch>();
        h->impl = std::make_unique<v2h::Batch>(ds, to_opts(opt), device);
        *out = h.release();
    });
}

int v2h_batch_destroy(v2h_batch* b) {
    return guarded([&] { delete b; });
}

int v2h_batch_count(const v2h_batch* b, int* count) {
    return guarded([&] {
        check_ptr(count);
        *count = B(b).count();
    });
}

int v2h_batch_set_options(v2h_batch* b, const v2h_sdp_options* opt) {
    return guarded([&] {
        check_ptr(opt);
        B(b).set_options(to_opts(opt));
    });
}

int v2h_batch_set_cg_tolerance_factor(v2h_batch* b, double factor) {
    return guarded([&] { B(b).set_cg_tolerance_factor(factor); });
}

int v2h_batch_sizes(const v2h_batch* b, int inst, int64_t* np, int64_t* nd, int* nb) {
    return guarded([&] {
        check_inst(B(b), inst);
        if (np) *np = B(b).n_primal(inst);
        if (nd) *nd = B(b).n_dual(inst);
        if (nb) *nb = (int)B(b).layout(inst).blocks.size();
    });
}

int v2h_batch_block_dims(const v2h_batch* b, int inst, int* dims) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(dims);
        const auto& bl = B(b).layout(inst).blocks;
        for (size_t k = 0; k < bl.size(); k++) dims[k] = bl[k].dim;
    });
}

int v2h_batch_orbital_order(const v2h_batch* b, int inst, int* to_user) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(to_user);
        const auto& v = B(b).layout(inst).to_user;
        for (size_t k = 0; k < v.size(); k++) to_user[k] = v[k];
    });
}

int v2h_batch_set_integrals(v2h_batch* b, int inst, const double* h, const double* g) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(h);
        check_ptr(g);
        B(b).set_integrals(inst, h, g);
    });
}

int v2h_batch_set_df_integrals(v2h_batch* b, int inst, int nfull, const double* h_full, int nq, const double* qmo,
                               int ncore, const int* core, const int* active, double* efzc) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(h_full);
        check_ptr(qmo);
        check_ptr(active);
        if (ncore > 0) check_ptr(core);
        const double e = B(b).set_df_integrals(inst, nfull, h_full, nq, qmo, ncore, core, active);
        if (efzc) *efzc = e;
    });
}

int v2h_batch_write_fcidump(v2h_batch* b, int inst, int nfull, const int* full_irreps, const double* h_full, int nq,
                            const double* qmo, int ncore, const int* core, const int* active, double enuc,
                            const char* int_path, const char* rdm_path, int64_t* n_int, int64_t* n_rdm) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(h_full);
        check_ptr(qmo);
        check_ptr(active);
        check_ptr(int_path);
        check_ptr(rdm_path);
        if (ncore > 0) check_ptr(core);
        v2h::FcidumpArgs a;
        a.nfull = nfull;
        a.full_irreps = full_irreps;
        a.h_full = h_full;
        a.nq = nq;
        a.qmo = qmo;
        a.ncore = ncore;
        a.core = core;
        a.active = active;
        a.enuc = enuc;
        a.int_path = int_path;
        a.rdm_path = rdm_path;
        B(b).write_fcidump(inst, a, n_int, n_rdm);
    });
}

int v2h_batch_optimize_orbitals(v2h_batch* b, int inst, const int* full_irreps, const v2h_orbopt_options* opt,
                                v2h_orbopt_result* res) {
    return guarded([&] {
        check_inst(B(b), inst);
        v2h::OrbOptArgs a;
        a.full_irreps = full_irreps;
        if (opt) {
            a.g_convergence = opt->g_convergence;
            a.e_convergence = opt->e_convergence;
            a.maxiter = opt->maxiter;
            a.active_active = opt->active_active;
        }
        v2h::OrbOptResult r;
        B(b).optimize_orbitals(inst, a, &r);
        if (res) {
            res->iterations = r.iterations;
            res->converged = r.converged;
            res->nrot = r.nrot;
            res->evaluations = r.evaluations;
            res->rotations = r.rotations;
            res->gnorm = r.gnorm;
            res->dE = r.dE;
            res->energy = r.energy;
            res->efzc = r.efzc;
        }
    });
}

int v2h_batch_orbital_gradient(v2h_batch* b, int inst, const int* full_irreps, int active_active, int32_t* pairs,
                               double* g, double* H, int cap, int* nrot) {
    return guarded([&] {
        check_inst(B(b), inst);
        v2h::OrbOptArgs a;
        a.full_irreps = full_irreps;
        a.active_active = active_active;
        const int n = B(b).orbital_gradient(inst, a, pairs, g, H, cap);
        if (nrot) *nrot = n;
    });
}

int v2h_batch_get_df_orbitals(v2h_batch* b, int inst, double* U, double* h_full, double* qmo) {
    return guarded([&] {
        check_inst(B(b), inst);
        B(b).get_df_orbitals(inst, U, h_full, qmo);
    });
}

int v2h_batch_set_cost(v2h_batch* b, int inst, const double* c) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(c);
        B(b).set_cost(inst, c);
    });
}

int v2h_batch_set_guess_hf(v2h_batch* b, int inst) {
    return guarded([&] {
        check_inst(B(b), inst);
        B(b).set_guess_hf(inst);
    });
}

int v2h_batch_set_guess(v2h_batch* b, int inst, const double* x) {
    return guarded([&] {
        check_inst(B(b), inst);
        B(b).set_guess(inst, x);
    });
}

int v2h_batch_reset(v2h_batch* b, int inst) {
    return guarded([&] {
        check_inst(B(b), inst);
        B(b).reset_state(inst);
    });
}

int v2h_batch_solve(v2h_batch* b, int maxiter, v2h_monitor_fn mon, void* data, int* n_converged) {
    return guarded([&] {
        if (maxiter < 0) throw std::invalid_argument("maxiter must be non-negative");
        v2h::Monitor m;
        if (mon)
            m = [&](int i, const v2h::InstStatus& s) {
                mon(i, (int)s.oiter, s.cg_last, s.energy_primal, s.energy_dual, s.mu, s.eps_p, s.eps_d, data);
            };
        int nc = B(b).solve(maxiter, m);
        if (n_converged) *n_converged = nc;
    });
}

int v2h_batch_status(v2h_batch* b, int inst, v2h_status* st) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(st);
        v2h::InstStatus s = B(b).status(inst);
        st->mu = s.mu;
        st->energy_primal = s.energy_primal;
        st->energy_dual = s.energy_dual;
        st->primal_error = s.eps_p;
        st->dual_error = s.eps_d;
        st->oiter_total = s.oiter;
        st->iiter_total = s.iiter;
        st->converged = s.converged;
        st->active = s.active;
        st->status_bits = s.status;
        st->last_cg_iterations = s.cg_last;
    });
}

double v2h_batch_last_solve_seconds(const v2h_batch* b) {
    return (b && b->impl) ? b->impl->last_solve_seconds() : -1.0;
}

#define V2H_VEC_GETTER(name, method)                              \
    int name(v2h_batch* b, int inst, double* out) {                \
        return guarded([&] {                                        \
            check_inst(B(b), inst);                                 \
            check_ptr(out);                                         \
            B(b).method(inst, out);                                 \
        });                                                         \
    }
V2H_VEC_GETTER(v2h_batch_get_x, get_primal)
V2H_VEC_GETTER(v2h_batch_get_y, get_dual)
V2H_VEC_GETTER(v2h_batch_get_z, get_slack)
V2H_VEC_GETTER(v2h_batch_get_c, get_cost)
V2H_VEC_GETTER(v2h_batch_get_b, get_b)
#undef V2H_VEC_GETTER

int v2h_batch_get_opdm(v2h_batch* b, int inst, const char* type, double* out) {
    return guarded([&] {
        check_inst(B(b), inst);
        if (!type || !out) throw std::invalid_argument("null argument");
        std::vector<double> x(B(b).n_primal(inst));
        B(b).get_primal(inst, x.data());
        v2h::opdm_dense(B(b).layout(inst), x.data(), type, out);
    });
}

int v2h_batch_get_tpdm(v2h_batch* b, int inst, double* out) {
    return guarded([&] {
        check_inst(B(b), inst);
        if (!out) throw std::invalid_argument("null argument");
        std::vector<double> x(B(b).n_primal(inst));
        B(b).get_primal(inst, x.data());
        v2h::tpdm_dense(B(b).layout(inst), x.data(), out);
    });
}

int v2h_batch_get_opdm_sparse(v2h_batch* b, int inst, const char* type, int32_t* idx, double* val, int64_t cap,
                              int64_t* count) {
    return guarded([&] {
        check_inst(B(b), inst);
        if (!type || !count || cap < 0 || (cap > 0 && (!idx || !val))) throw std::invalid_argument("bad argument");
        std::vector<double> x(B(b).n_primal(inst));
        B(b).get_primal(inst, x.data());
        auto e = v2h::opdm_sparse(B(b).layout(inst), x.data(), type);
        *count = (int64_t)e.size();
        for (int64_t k = 0; k < cap && k < (int64_t)e.size(); k++) {
            idx[2 * k] = e[k].i;
            idx[2 * k + 1] = e[k].j;
            val[k] = e[k].value;
        }
    });
}

int v2h_batch_get_tpdm_sparse(v2h_batch* b, int inst, const char* type, int32_t* idx, double* val, int64_t cap,
                              int64_t* count) {
    return guarded([&] {
        check_inst(B(b), inst);
        if (!type || !count || cap < 0 || (cap > 0 && (!idx || !val))) throw std::invalid_argument("bad argument");
        std::vector<double> x(B(b).n_primal(inst));
        B(b).get_primal(inst, x.data());
        auto e = v2h::tpdm_sparse(B(b).layout(inst), x.data(), type);
        *count = (int64_t)e.size();
        for (int64_t k = 0; k < cap && k < (int64_t)e.size(); k++) {
            idx[4 * k] = e[k].i;
            idx[4 * k + 1] = e[k].j;
            idx[4 * k + 2] = e[k].k;
            idx[4 * k + 3] = e[k].l;
            val[k] = e[k].value;
        }
    });
}

int v2h_batch_set_y(v2h_batch* b, int inst, const double* y) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(y);
        B(b).set_dual(inst, y);
    });
}

int v2h_batch_set_z(v2h_batch* b, int inst, const double* z) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(z);
        B(b).set_slack(inst, z);
    });
}

int v2h_batch_get_mu(v2h_batch* b, int inst, double* mu) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(mu);
        *mu = B(b).get_mu(inst);
    });
}

int v2h_batch_set_mu(v2h_batch* b, int inst, double mu) {
    return guarded([&] {
        check_inst(B(b), inst);
        if (!(mu > 0.0)) throw std::invalid_argument("mu must be positive");
        B(b).set_mu(inst, mu);
    });
}

int v2h_batch_Au(v2h_batch* b, int inst, const double* u, double* out) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(u);
        check_ptr(out);
        B(b).apply_A(inst, u, out);
    });
}

namespace {
v2h::FullSpace full_space(const v2h::Layout& L, int ncore, const int* core, const int* active) {
    if (ncore < 0) throw std::invalid_argument("ncore must be non-negative");
    if (ncore > 0 && !core) throw std::invalid_argument("null core list");
    if (!active) throw std::invalid_argument("null active list");
    v2h::FullSpace f;
    f.core.assign(core, core + ncore);
    f.active.assign(active, active + L.n);
    return f;
}
}  // namespace

int v2h_batch_write_rdm_files(v2h_batch* b, int inst, int ncore, const int* core, const int* active,
                              const char* prefix, int which, int64_t* counts) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(prefix);
        check_ptr(counts);
        const v2h::Layout& L = B(b).layout(inst);
        std::vector<double> x(B(b).n_primal(inst));
        B(b).get_primal(inst, x.data());
        v2h::write_rdm_files(L, x.data(), full_space(L, ncore, core, active), prefix, which, counts);
    });
}

int v2h_batch_Au_noalloc(v2h_batch* b, int inst, const double* u, double* out) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(u);
        check_ptr(out);
        v2h::Batch& bt = B(b);
        const int64_t np = bt.n_primal(inst), nd = bt.n_dual(inst);
        if (np + nd > v2h::kNoAllocDoubles)
            throw std::invalid_argument("v2h_batch_Au_noalloc: n_primal + n_dual = " + std::to_string(np + nd) +
                                        " exceeds the scratch of " + std::to_string(v2h::kNoAllocDoubles) +
                                        " doubles");
        V2H_HIP(hipSetDevice(bt.device()));
        double* s = v2h::noalloc_scratch(bt.device());  // the scratch's device address, never its name
        hipStream_t st = bt.stream();
        V2H_HIP(hipMemcpyAsync(s, u, sizeof(double) * (size_t)np, hipMemcpyHostToDevice, st));
        bt.apply_A_device(inst, s, s + np);
        V2H_HIP(hipMemcpyAsync(out, s + np, sizeof(double) * (size_t)nd, hipMemcpyDeviceToHost, st));
        V2H_HIP(hipStreamSynchronize(st));
    });
}

int v2h_batch_ATu(v2h_batch* b, int inst, const double* y, double* out) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(y);
        check_ptr(out);
        B(b).apply_AT(inst, y, out);
    });
}

int v2h_batch_set_profiling(v2h_batch* b, int on) {
    return guarded([&] { B(b).set_profiling(on != 0); });
}

static_assert(sizeof(v2h_tuning) == sizeof(int) * (2 * 7 + 5), "v2h_tuning layout");

int v2h_batch_get_tuning(const v2h_batch* b, v2h_tuning* t) {
    return guarded([&] {
        check_ptr(t);
        const auto& u = B(const_cast<v2h_batch*>(b)).tuning();
        for (int k = 0; k < 7; k++) {
            t->class_lane[k] = u.class_lane[k];
            t->fallback_lane[k] = u.fallback_lane[k];
        }
        t->refine_lane = u.refine_lane;
        t->cg_margin = u.cg_margin;
        t->xcd_remap = u.xcd_remap;
        t->graphs = u.graphs;
        t->pipeline = u.pipeline;
    });
}

int v2h_batch_set_tuning(v2h_batch* b, const v2h_tuning* t) {
    return guarded([&] {
        check_ptr(t);
        v2h::Batch::Tuning u;
        for (int k = 0; k < 7; k++) {
            u.class_lane[k] = t->class_lane[k];
            u.fallback_lane[k] = t->fallback_lane[k];
        }
        u.refine_lane = t->refine_lane;
        u.cg_margin = t->cg_margin;
        u.xcd_remap = t->xcd_remap;
        u.graphs = t->graphs;
        u.pipeline = t->pipeline;
        B(b).set_tuning(u);
    });
}

int v2h_batch_set_graphs(v2h_batch* b, int mode) {
    return guarded([&] {
        V2H_HIP(hipSetDevice(B(b).device()));
        B(b).set_graphs(mode);
    });
}

int v2h_batch_kernel_times(v2h_batch* b, double* ms, int64_t* counts, double* full_ms, int64_t* full_counts) {
    return guarded([&] { B(b).kernel_times(ms, counts, full_ms, full_counts); });
}

int v2h_batch_reset_kernel_times(v2h_batch* b) {
    return guarded([&] { B(b).reset_kernel_times(); });
}

int v2h_batch_launch_stats(v2h_batch* b, int64_t* st) {
    return guarded([&] {
        check_ptr(st);
        B(b).launch_stats(st);
    });
}

int v2h_batch_set_projection(v2h_batch* b, int mode) {
    return guarded([&] {
        if (mode != V2H_PROJECT_TRIDIAGONAL && mode != V2H_PROJECT_REFINED)
            throw std::invalid_argument("unknown projection mode");
        B(b).set_refine(mode == V2H_PROJECT_REFINED);
    });
}

int v2h_batch_projection_stats(v2h_batch* b, int64_t* st) {
    return guarded([&] {
        check_ptr(st);
        B(b).projection_stats(st);
    });
}

int v2h_batch_set_refine_params(v2h_batch* b, double cluster_rel, double accept_e, double accept_cpl,
                                int max_steps) {
    return guarded([&] {
        if (!(cluster_rel >= 0.0 && cluster_rel < 1e-2) || !(accept_e > 0.0 && accept_e <= 1e-3) ||
            !(accept_cpl > 0.0 && accept_cpl <= 1e-6) || max_steps < 1 || max_steps > v2h::kMaxSteps)
            throw std::invalid_argument("refinement thresholds out of range");
        v2h::RefParams p;
        p.cluster_rel = cluster_rel;
        p.accept_e = accept_e;
        p.accept_cpl = accept_cpl;
        p.max_steps = max_steps;
        B(b).set_refine_params(p);
    });
}

int v2h_batch_cost_adjoint(v2h_batch* b, int inst, const double* xbar, double* dh, double* dg) {
    return guarded([&] {
        check_inst(B(b), inst);
        check_ptr(dh);
        check_ptr(dg);
        B(b).cost_adjoint(inst, xbar, dh, dg);
    });
}

}  // extern "C"

// ---- host-only structure inspection -------------------------------------

struct v2h_sdp {
    std::unique_ptr<v2h::CallbackSdp> impl;
};

namespace {
v2h::CallbackSdp& S(v2h_sdp* s) {
    if (!s || !s->impl) throw std::invalid_argument("null solver handle");
    return *s->impl;
}
const v2h::CallbackSdp& S(const v2h_sdp* s) {
    if (!s || !s->impl) throw std::invalid_argument("null solver handle");
    return *s->impl;
}
}  // namespace

extern "C" {

int v2h_sdp_create(int64_t n_primal, int64_t n_dual, const v2h_sdp_options* opt, int device, v2h_sdp** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output handle");
        *out = nullptr;
        v2h_sdp_options o;
        v2h_default_options(&o);
        auto h = std::make_unique<v2h_sdp>();
        h->impl = std::make_unique<v2h::CallbackSdp>(n_primal, n_dual, to_opts(opt ? opt : &o), device);
        *out = h.release();
    });
}

int v2h_sdp_destroy(v2h_sdp* s) {
    return guarded([&] { delete s; });
}

int v2h_sdp_set_options(v2h_sdp* s, const v2h_sdp_options* opt) {
    return guarded([&] {
        if (!opt) throw std::invalid_argument("null options");
        S(s).set_options(to_opts(opt));
    });
}

int v2h_sdp_solve(v2h_sdp* s, double* x, const double* b, const double* c, const int* block_dims, int n_blocks,
                  int maxiter, v2h_map_fn Au, v2h_map_fn ATu, v2h_sdp_monitor_fn mon, int print_level, void* data,
                  int* converged) {
    return guarded([&] {
        int cv = S(s).solve(x, b, c, block_dims, n_blocks, maxiter, Au, ATu, mon, print_level, data);
        if (converged) *converged = cv;
    });
}

int v2h_sdp_status(const v2h_sdp* s, v2h_status* st) {
    return guarded([&] {
        if (!st) throw std::invalid_argument("null status");
        const auto& q = S(s);
        st->mu = q.mu();
        st->energy_primal = q.energy_primal;
        st->energy_dual = q.energy_dual;
        st->primal_error = q.eps_p;
        st->dual_error = q.eps_d;
        st->oiter_total = q.oiter;
        st->iiter_total = q.iiter;
        st->converged = q.converged;
        st->active = 0;
        st->status_bits = q.status;
        st->last_cg_iterations = q.last_cg;
    });
}

int v2h_sdp_get_y(const v2h_sdp* s, double* y) {
    return guarded([&] {
        if (!y) throw std::invalid_argument("null argument");
        std::copy(S(s).y().begin(), S(s).y().end(), y);
    });
}
int v2h_sdp_get_z(const v2h_sdp* s, double* z) {
    return guarded([&] {
        if (!z) throw std::invalid_argument("null argument");
        std::copy(S(s).z().begin(), S(s).z().end(), z);
    });
}
int v2h_sdp_set_y(v2h_sdp* s, const double* y) {
    return guarded([&] {
        if (!y) throw std::invalid_argument("null argument");
        S(s).set_y(y);
    });
}
int v2h_sdp_set_z(v2h_sdp* s, const double* z) {
    return guarded([&] {
        if (!z) throw std::invalid_argument("null argument");
        S(s).set_z(z);
    });
}
int v2h_sdp_get_mu(const v2h_sdp* s, double* mu) {
    return guarded([&] {
        if (!mu) throw std::invalid_argument("null argument");
        *mu = S(s).mu();
    });
}
int v2h_sdp_set_mu(v2h_sdp* s, double mu) {
    return guarded([&] { S(s).set_mu(mu); });
}

}  // extern "C"

struct v2h_structure {
    v2h::Layout L;
    v2h::Csr A;
    v2h::CostMap cost;
    v2h::SellHost sA, sAT;
    explicit v2h_structure(const v2h::ProblemDesc& d) : L(d) {
        A = v2h::build_constraints(L);
        cost = v2h::build_cost_map(L);
        sA = v2h::build_sell(A, 32);
        sAT = v2h::build_sell(A.transpose(), 32);
    }
};

namespace {
const v2h_structure& Sref(const v2h_structure* s) {
    if (!s) throw std::invalid_argument("null structure handle");
    return *s;
}
}  // namespace

extern "C" {

int v2h_structure_create(const v2h_problem_desc* s, v2h_structure** out) {
    return guarded([&] {
        check_ptr(s);
        check_ptr(out);
        *out = nullptr;
        v2h::ProblemDesc d;
        d.nmo = s->nmo;
        d.nalpha = s->nalpha;
        d.nbeta = s->nbeta;
        d.multiplicity = s->multiplicity;
        d.nirrep = s->nirrep;
        if (s->orbital_irreps) d.irreps.assign(s->orbital_irreps, s->orbital_irreps + s->nmo);
        d.q2 = s->constrain_q2 != 0;
        d.g2 = s->constrain_g2 != 0;
        d.spin = s->constrain_spin != 0;
        d.sz = s->constrain_sz != 0;
        d.t1 = s->constrain_t1 != 0;
        d.t2 = s->constrain_t2 != 0;
        d.d3 = s->constrain_d3 != 0;
        *out = new v2h_structure(d);
    });
}

int v2h_structure_destroy(v2h_structure* s) {
    return guarded([&] { delete s; });
}

int v2h_structure_sizes(const v2h_structure* s, int64_t* np, int64_t* nd, int64_t* nnz, int* nb) {
    return guarded([&] {
        const auto& S = Sref(s);
        if (np) *np = S.L.n_primal;
        if (nd) *nd = S.A.nrows;
        if (nnz) *nnz = S.A.nnz();
        if (nb) *nb = (int)S.L.blocks.size();
    });
}

int v2h_structure_block_dims(const v2h_structure* s, int* dims) {
    return guarded([&] {
        check_ptr(dims);
        const auto& bl = Sref(s).L.blocks;
        for (size_t k = 0; k < bl.size(); k++) dims[k] = bl[k].dim;
    });
}

int v2h_structure_tables(const v2h_structure* s, int* gab, int* gaa, int* bab, int* baa) {
    return guarded([&] {
        const auto& L = Sref(s).L;
        int kab = 0, kaa = 0;
        for (int h = 0; h < L.H; h++) {
            if (gab) gab[h] = L.gab[h];
            if (gaa) gaa[h] = L.gaa[h];
            for (const auto& p : L.ab_pairs[h]) {
                if (bab) { bab[kab] = p[0]; bab[kab + 1] = p[1]; }
                kab += 2;
            }
            for (const auto& p : L.aa_pairs[h]) {
                if (baa) { baa[kaa] = p[0]; baa[kaa + 1] = p[1]; }
                kaa += 2;
            }
        }
    });
}

int v2h_structure_constraints(const v2h_structure* s, int64_t* ptr, int32_t* col, double* val, double* b) {
    return guarded([&] {
        const auto& A = Sref(s).A;
        if (ptr) std::copy(A.ptr.begin(), A.ptr.end(), ptr);
        if (col) std::copy(A.col.begin(), A.col.end(), col);
        if (val) std::copy(A.val.begin(), A.val.end(), val);
        if (b) std::copy(A.rhs.begin(), A.rhs.end(), b);
    });
}

int v2h_structure_cost_map(const v2h_structure* s, int32_t* i1, int32_t* i2) {
    return guarded([&] {
        const auto& c = Sref(s).cost;
        if (i1) std::copy(c.i1.begin(), c.i1.end(), i1);
        if (i2) std::copy(c.i2.begin(), c.i2.end(), i2);
    });
}

int v2h_structure_gather_stats(const v2h_structure* s, int64_t* st) {
    return guarded([&] {
        check_ptr(st);
        const auto& S = Sref(s);
        st[0] = S.sA.nslices;
        st[1] = (int64_t)S.sA.ent.size();
        st[2] = S.sA.nlong();
        st[3] = (int64_t)S.sA.long_ent.size();
        st[4] = S.sAT.nslices;
        st[5] = (int64_t)S.sAT.ent.size();
        st[6] = S.sAT.nlong();
        st[7] = (int64_t)S.sAT.long_ent.size();
    });
}

int v2h_structure_opdm(const v2h_structure* s, const double* x, const char* type, double* out) {
    return guarded([&] {
        check_ptr(x);
        check_ptr(type);
        check_ptr(out);
        v2h::opdm_dense(Sref(s).L, x, type, out);
    });
}

int v2h_structure_write_rdm_files(const v2h_structure* s, const double* x, int ncore, const int* core,
                                  const int* active, const char* prefix, int which, int64_t* counts) {
    return guarded([&] {
        check_ptr(x);
        check_ptr(prefix);
        check_ptr(counts);
        const v2h::Layout& L = Sref(s).L;
        v2h::write_rdm_files(L, x, full_space(L, ncore, core, active), prefix, which, counts);
    });
}

int v2h_structure_hf_guess(const v2h_structure* s, double* x) {
    return guarded([&] {
        check_ptr(x);
        const auto& S = Sref(s);
        std::vector<double> g = v2h::hf_guess(S.L, S.A, S.L.na, S.L.na);
        std::copy(g.begin(), g.end(), x);
    });
}

int v2h_structure_tpdm(const v2h_structure* s, const double* x, double* out) {
    return guarded([&] {
        check_ptr(x);
        check_ptr(out);
        v2h::tpdm_dense(Sref(s).L, x, out);
    });
}

}  // extern "C"
