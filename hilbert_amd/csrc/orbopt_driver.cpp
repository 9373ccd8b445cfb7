// orbopt_driver.cpp -- one call of the v2RDM-CASSCF orbital optimiser on the device
// for one instance of a batch.
//
// Mirrors v2RDMSolver::RotateOrbitals (v2rdm_casscf/v2rdm_solver.cc:3221-3292)
// with the density-fitted optimiser OrbitalOptimizer::optimize_orbitals
// (focas/orbopt_optimize.cc:25-110): the densities of the current primal
// vector stay fixed for the call; each iteration that evaluates the gradient
// forms the generalized Fock matrix, the gradient and the exact diagonal
// Hessian, takes the preconditioned step kappa = -r g / max(|H|, 1e-2),
// rotates the resident integrals by exp(K(kappa)) and re-runs the DF intake
// (RepackIntegrals, tei.cc:139-199) to get the new energy c.x + E_fzc; the
// trust factor r grows by 1.2 when the energy falls by more than 3/4 of the
// quadratic model, shrinks by 0.7 below 1/4 or when the energy rises, and a
// rise replays the step scaled down (Update_r, Update_kappa).  A final rise
// is undone (Restore_last_step).  Everything heavy runs on the device; the
// host reads back three scalars per iteration (|g|^2, the model change and
// the energy) and the 1-norm of K for the exponential's scaling.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "orbopt.hpp"
#include "solver.hpp"

namespace v2h {

void Batch::oo_setup(int i, const OrbOptArgs& a) {
    DfState& S = df_state(i);
    const Layout& L = layout(i);
    const int nfull = S.nfull, n = S.n, ncore = S.ncore, no = S.no, nq = S.nq;
    if (a.maxiter < 1) throw std::invalid_argument("optimize_orbitals: maxiter must be >= 1");
    if (!(a.g_convergence >= 0.0) || !(a.e_convergence >= 0.0))
        throw std::invalid_argument("optimize_orbitals: convergence thresholds must be >= 0");
    auto irr = [&](int p) { return a.full_irreps ? a.full_irreps[p] : 0; };
    for (int p = 0; p < nfull; p++)
        if (irr(p) < 0 || irr(p) >= kMaxIrrep)
            throw std::invalid_argument("optimize_orbitals: irrep label out of range");
    for (int u = 0; u < n; u++) {
        const int pz = (int)(std::find(L.to_user.begin(), L.to_user.end(), u) - L.to_user.begin());
        if (irr(S.active[u]) != L.sym[pz])
            throw std::invalid_argument("optimize_orbitals: active orbital irreps differ from the instance's");
    }
    std::vector<int> key(nfull + 1);
    for (int p = 0; p < nfull; p++) key[p] = irr(p);
    key[nfull] = a.active_active ? 1 : 0;
    if (S.nrot >= 0 && key == S.pair_key) return;
    S.pair_key = key;

    // rotation pairs (p of the higher class), p-major, q ascending
    std::vector<int> cls(nfull, 2);
    for (int c : S.core) cls[c] = 0;
    for (int t : S.active) cls[t] = 1;
    std::vector<int2> pairs;
    for (int p = 0; p < nfull; p++)
        for (int q = 0; q < nfull; q++) {
            if (p == q || irr(p) != irr(q)) continue;
            if (cls[p] > cls[q] || (a.active_active && cls[p] == 1 && cls[q] == 1 && p > q)) pairs.push_back(make_int2(p, q));
        }
    S.nrot = (int)pairs.size();
    if (pairs.empty()) pairs.push_back(make_int2(0, 0));  // keeps the buffer non-empty; never read
    S.pairs.upload(pairs);

    // O = core (given order) + active (Pitzer order)
    std::vector<int> occ(no), pos(nfull, -1);
    for (int c = 0; c < ncore; c++) occ[c] = S.core[c];
    for (int k = 0; k < n; k++) occ[ncore + k] = S.active[L.to_user[k]];
    for (int o = 0; o < no; o++) pos[occ[o]] = o;
    S.occ.upload(occ);
    S.pos.upload(pos);

    std::vector<int4> ptab((size_t)n * n);
    std::vector<int64_t> d1((size_t)2 * n * n, -1);
    for (int pi = 0; pi < n; pi++)
        for (int pj = 0; pj < n; pj++) {
            const int h = L.pair_irrep(pi, pj);
            ptab[(size_t)pi * n + pj] = make_int4(h, L.ab(h, pi, pj), pi != pj ? L.aa(h, pi, pj) : -1, pi > pj ? -1 : 1);
            if (L.sym[pi] == L.sym[pj]) {
                d1[(size_t)pi * n + pj] = L.one_body(D1A, pi, pj);
                d1[(size_t)n * n + pi * n + pj] = L.one_body(D1B, pi, pj);
            }
        }
    S.ptab.upload(ptab);
    S.d1tab.upload(d1);
    OoTables& t = S.tabs;
    t = OoTables{};
    t.n = n;
    t.ncore = ncore;
    t.no = no;
    t.pair = S.ptab.p;
    t.d1 = S.d1tab.p;
    for (int h = 0; h < kMaxIrrep; h++) {
        if (h < L.H) {
            t.offab[h] = L.off[D2AB][h];
            t.offaa[h] = L.off[D2AA][h];
            t.offbb[h] = L.off[D2BB][h];
            t.gab[h] = L.gab[h];
            t.gaa[h] = L.gaa[h];
        }
    }

    const size_t nn = (size_t)nfull * nfull, n2 = (size_t)no * no, nr = std::max(1, S.nrot);
    if (!S.bd.p) {
        S.bd.alloc((size_t)nq * nn);
        S.tmp.alloc((size_t)nq * nn);
        for (auto* b : {&S.uscr, &S.hscr, &S.K, &S.U, &S.T1}) b->alloc(nn);
        S.D.alloc(n2);
        S.d.alloc(n2 * n2);
        S.gO.alloc(n2 * n2);
        S.BO.alloc((size_t)nq * n2);
        S.BFO.alloc((size_t)nfull * no * nq);
        S.Bdiag.alloc((size_t)nfull * nq);
        S.Z.alloc(n2 * nq);
        S.F.alloc((size_t)no * nfull);
        S.Fw.alloc((size_t)kMaxSlices * no * nfull);
        S.Jw.alloc((size_t)dgemm_split_slices(nfull, (int)n2, nq) * nfull * n2);
        S.Gw.alloc((size_t)dgemm_split_slices((int)n2, (int)n2, nq) * n2 * n2);
        S.Zw.alloc((size_t)dgemm_split_slices((int)n2, nq, (int)n2) * n2 * nq);
        S.Jd.alloc((size_t)nfull * n2);
        S.Kd.alloc((size_t)nfull * n2);
        S.red.alloc(3);
        if (!S.pin) V2H_HIP(hipHostMalloc((void**)&S.pin, 8 * sizeof(double), hipHostMallocDefault));
    }
    for (auto* b : {&S.g, &S.H, &S.kref, &S.kcur, &S.ksave}) b->alloc(nr);
}

void Batch::oo_gradient(int i) {
    DfState& S = df_state(i);
    const int nfull = S.nfull, no = S.no, nq = S.nq, n2 = S.no * S.no;
    tstart(kKindOrb);
    launch_oo_gather(S.bd.p, nq, nfull, S.occ.p, no, S.BO.p, S.BFO.p, S.Bdiag.p, stream_);
    launch_dgemm_split(true, false, n2, n2, nq, 1.0, S.BO.p, n2, S.BO.p, n2, 0.0, S.gO.p, n2, S.Gw.p, stream_);
    launch_dgemm_split(false, true, n2, nq, n2, 1.0, S.d.p, n2, S.BO.p, n2, 0.0, S.Z.p, nq, S.Zw.p, stream_);
    launch_dgemm_split(false, true, no, nfull, no * nq, 1.0, S.Z.p, (int64_t)no * nq, S.BFO.p, (int64_t)no * nq, 0.0,
                       S.F.p, nfull, S.Fw.p, stream_);
    launch_oo_fock1(S.D.p, S.h.p, S.occ.p, no, nfull, S.F.p, stream_);
    launch_dgemm_split(false, false, nfull, n2, nq, 1.0, S.Bdiag.p, nq, S.BO.p, n2, 0.0, S.Jd.p, n2, S.Jw.p, stream_);
    launch_dgemm(false, true, no, no, nq, 1.0, S.BFO.p, nq, (int64_t)no * nq, S.BFO.p, nq, (int64_t)no * nq, 0.0,
                 S.Kd.p, no, n2, nfull, stream_);
    OoGradArgs ga{};
    ga.pairs = S.pairs.p;
    ga.nrot = S.nrot;
    ga.nfull = nfull;
    ga.no = no;
    ga.pos = S.pos.p;
    ga.F = S.F.p;
    ga.D = S.D.p;
    ga.d = S.d.p;
    ga.h = S.h.p;
    ga.gO = S.gO.p;
    ga.Jd = S.Jd.p;
    ga.Kd = S.Kd.p;
    ga.g = S.g.p;
    ga.H = S.H.p;
    launch_oo_grad_hess(ga, stream_);
    tstop(kKindOrb);
    V2H_HIP(hipGetLastError());
}

double Batch::oo_apply(int i, const double* kappa, double sign) {
    DfState& S = df_state(i);
    const int nfull = S.nfull, nq = S.nq;
    const int64_t nn = (int64_t)nfull * nfull;
    tstart(kKindOrb);
    launch_oo_kappa_matrix(kappa, S.pairs.p, S.nrot, sign, nfull, S.K.p, S.red.p + 2, stream_);
    tstop(kKindOrb);
    // one sync: ||K||_1, and with it the step model copied after the last
    // preconditioning
    V2H_HIP(hipMemcpyAsync(S.pin + 2, S.red.p + 2, sizeof(double), hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
    const double nrm = S.pin[2];
    if (!std::isfinite(nrm)) throw std::runtime_error("optimize_orbitals: non-finite rotation");
    int sq = 0;
    while (sq < 60 && nrm / std::ldexp(1.0, sq) > 0.5) sq++;
    tstart(kKindOrb);
    if (sq > 0) launch_oo_axpy(std::ldexp(1.0, -sq) - 1.0, S.K.p, S.K.p, nn, stream_);
    // exp(K) on the scaled K: the Taylor polynomial of degree 14 by Horner's
    // rule, T <- I + K T / k for k = 14 .. 1 (the identity added in the
    // product's epilogue), then squared back
    double* U = S.U.p;
    double* T1 = S.T1.p;
    launch_oo_identity(U, nfull, stream_);
    constexpr int kTaylor = 14;
    for (int k = kTaylor; k >= 1; k--) {
        launch_dgemm_small(false, nfull, nfull, nfull, 1.0 / k, S.K.p, nfull, U, nfull, T1, nfull, stream_, 1.0);
        std::swap(U, T1);
    }
    for (int k = 0; k < sq; k++) {
        launch_dgemm_small(false, nfull, nfull, nfull, 1.0, U, nfull, U, nfull, T1, nfull, stream_);
        std::swap(U, T1);
    }
    // B^Q <- U^T B^Q U, h <- U^T h U, Uacc <- Uacc U
    launch_dgemm(false, false, nfull, nfull, nfull, 1.0, S.bd.p, nfull, nn, U, nfull, 0, 0.0, S.tmp.p, nfull, nn, nq,
                 stream_);
    launch_dgemm(true, false, nfull, nfull, nfull, 1.0, U, nfull, 0, S.tmp.p, nfull, nn, 0.0, S.bd.p, nfull, nn, nq,
                 stream_);
    launch_dgemm_small(false, nfull, nfull, nfull, 1.0, S.h.p, nfull, U, nfull, S.hscr.p, nfull, stream_);
    launch_dgemm_small(true, nfull, nfull, nfull, 1.0, U, nfull, S.hscr.p, nfull, S.h.p, nfull, stream_);
    launch_dgemm_small(false, nfull, nfull, nfull, 1.0, S.uacc.p, nfull, U, nfull, S.uscr.p, nfull, stream_);
    std::swap(S.uacc.p, S.uscr.p);
    launch_df_pack(S.bd.p, nq, nfull, S.qmo.p, stream_);
    tstop(kKindOrb);
    V2H_HIP(hipGetLastError());
    // the intake (||c|| once, at the end of the call) and E = c.x + E_fzc,
    // read back together
    launch_df_intake_only(i);
    const InstDev& I = inst_[i];
    launch_dot1(c_.p + I.poff, x_.p + I.poff, n_primal(i), scratch_.p, stream_);
    V2H_HIP(hipMemcpyAsync(S.pin + 3, S.e.p, sizeof(double), hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipMemcpyAsync(S.pin + 4, scratch_.p, sizeof(double), hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
    S.efzc = S.pin[3];
    return S.pin[4] + S.efzc;
}

void Batch::optimize_orbitals(int i, const OrbOptArgs& a, OrbOptResult* res) {
    V2H_HIP(hipSetDevice(device_));
    DfState& S = df_state(i);
    oo_setup(i, a);
    const InstDev& I = inst_[i];
    const int nr = std::max(1, S.nrot);
    launch_df_unpack(S.qmo.p, S.nq, S.nfull, S.bd.p, stream_);
    launch_oo_density(S.tabs, x_.p + I.poff, S.D.p, S.d.p, stream_);
    for (auto* b : {&S.kref, &S.kcur, &S.ksave}) launch_fill(b->p, nr, 0.0, stream_);
    run_df_intake(i, false);
    double E_cur = instance_energy(i);

    bool eval_G = true, conv_e = false, conv_g = false;
    double r = 1.0, E_init = E_cur, initial_E = E_cur, dE_quad = 0.0, dE = 0.0, gnorm = 0.0;
    int iters = 0, evals = 0, applies = 0;
    for (int it = 0; it < a.maxiter; it++) {
        if (eval_G) {
            evals++;
            oo_gradient(i);
            launch_oo_precondition(S.g.p, S.H.p, S.nrot, r, S.kref.p, S.kcur.p, S.ksave.p, S.red.p, stream_);
            // read back with the next sync (in oo_apply, before anything uses them)
            V2H_HIP(hipMemcpyAsync(S.pin, S.red.p, 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
            E_init = E_cur;
            if (it == 0) initial_E = E_init;
        }
        E_cur = oo_apply(i, S.kcur.p, 1.0);
        if (eval_G) {
            gnorm = std::sqrt(S.pin[0]);
            dE_quad = S.pin[1];
        }
        applies++;
        dE = E_cur - E_init;
        if (gnorm < a.g_convergence) conv_g = true;
        if (std::fabs(dE) < a.e_convergence) conv_e = true;
        iters++;
        if (conv_e || conv_g || it == a.maxiter - 1) break;
        // Update_r (orbopt_optimize.cc:124-152)
        eval_G = !(dE > 0.0);
        const double ratio = dE / dE_quad;
        if (dE < 0.0) {
            if (ratio > 0.75) r *= 1.2;
            if (ratio < 0.25) r *= 0.7;
        } else {
            r *= 0.7;
        }
        if (!eval_G) launch_oo_update_kappa(S.kref.p, S.nrot, r, S.kcur.p, S.ksave.p, stream_);
    }
    if (dE > 0.0) {
        if (conv_e || conv_g) {
            E_cur = oo_apply(i, S.kcur.p, -1.0);
            applies++;
        }
        if (!eval_G) {
            E_cur = oo_apply(i, S.ksave.p, -1.0);
            applies++;
        }
    }
    // The call-level test of RotateOrbitals (v2rdm_solver.cc:3271-3274): the
    // net energy change of the whole call and the last gradient norm, not the
    // optimiser's per-step flags.
    const bool call_converged = std::fabs(E_cur - initial_E) < a.e_convergence && gnorm < a.g_convergence;
    // the cost vector changed: its norm, and unless the call converged the
    // SDP is not converged for it yet
    refresh_norms(i);
    if (!call_converged) {
        sync_state_to_host();
        state_[i].converged = 0;
        sync_state_to_device();
    }
    if (res) {
        res->iterations = iters;
        res->converged = call_converged ? 1 : 0;
        res->nrot = S.nrot;
        res->evaluations = evals;
        res->rotations = applies;
        res->gnorm = gnorm;
        res->dE = E_cur - initial_E;
        res->energy = E_cur;
        res->efzc = S.efzc;
    }
}

int Batch::orbital_gradient(int i, const OrbOptArgs& a, int* pairs, double* g, double* H, int cap) {
    V2H_HIP(hipSetDevice(device_));
    DfState& S = df_state(i);
    oo_setup(i, a);
    const InstDev& I = inst_[i];
    if (cap < S.nrot) return S.nrot;
    launch_df_unpack(S.qmo.p, S.nq, S.nfull, S.bd.p, stream_);
    launch_oo_density(S.tabs, x_.p + I.poff, S.D.p, S.d.p, stream_);
    oo_gradient(i);
    if (S.nrot > 0) {
        if (pairs) V2H_HIP(hipMemcpyAsync(pairs, S.pairs.p, sizeof(int2) * S.nrot, hipMemcpyDeviceToHost, stream_));
        if (g) V2H_HIP(hipMemcpyAsync(g, S.g.p, sizeof(double) * S.nrot, hipMemcpyDeviceToHost, stream_));
        if (H) V2H_HIP(hipMemcpyAsync(H, S.H.p, sizeof(double) * S.nrot, hipMemcpyDeviceToHost, stream_));
    }
    V2H_HIP(hipStreamSynchronize(stream_));
    return S.nrot;
}

void Batch::get_df_orbitals(int i, double* U, double* h_full, double* qmo) {
    V2H_HIP(hipSetDevice(device_));
    DfState& S = df_state(i);
    const int64_t nn = (int64_t)S.nfull * S.nfull, ntri = (int64_t)S.nfull * (S.nfull + 1) / 2;
    if (U) V2H_HIP(hipMemcpyAsync(U, S.uacc.p, sizeof(double) * nn, hipMemcpyDeviceToHost, stream_));
    if (h_full) V2H_HIP(hipMemcpyAsync(h_full, S.h.p, sizeof(double) * nn, hipMemcpyDeviceToHost, stream_));
    if (qmo) V2H_HIP(hipMemcpyAsync(qmo, S.qmo.p, sizeof(double) * S.nq * ntri, hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
}

}  // namespace v2h
