"""CPU restatement of the v2RDM-CASSCF orbital step (test infrastructure only).

Only tests, the smoke check and the CPU baseline use this module, and only
to check the device path; the product never calls it.

What it restates:

* ``PackSpatialDensity`` (v2rdm_casscf/v2rdm_solver.cc:3110-3219): the
  spin-free 1- and 2-RDM of the active space from the primal vector, here
  dense, Pitzer order, d_pqrs = sum_st <a+_ps a+_rt a_st a_qs>.
* ``OrbitalOptimizer::optimize_orbitals`` (focas/orbopt_optimize.cc:25-110)
  with its step control ``Precondition_Step`` (:216-235), ``Update_r``
  (:124-152), ``Update_kappa`` (:111-122), ``Compute_Quadratic_dE``
  (:163-177) and ``Restore_last_step`` (:195-214); gradient norm as in
  ``ComputeG`` (focas/orbopt_gradient.cc:26-44).
* The orbital gradient (orbopt_gradient.cc:46-146) and the exact diagonal
  Hessian (orbopt_hessian.cc:25-1048; Chaban, Schmidt and Gordon, Theor.
  Chem. Acc. 97, 88 (1997), Eqs. 4.6) written once for every rotation class
  through the generalized Fock matrix and the exact second derivative along
  one rotation (Helgaker, Joergensen and Olsen, Molecular Electronic-Structure
  Theory, Sec. 10.8), over the occupied space O = core + active with the
  closed-shell core folded into D and d.
* ``ExponentiateKappa`` (orbopt_exponential.cc:25-187; U = exp(K)) and
  ``TransformIntegrals`` (orbopt_integraltransform_cpu.cc:25-262: h and every
  DF slice B^Q go to U^T X U, the orbital matrix to C U).

Convention here: rotation pair (p, q) (p of the higher class: virtual >
active > core, or p > q for active-active) has K[p, q] = kappa, K[q, p] =
-kappa; new orbital r = sum_s old_s U[s, r].  The gradient is then
dE/dkappa = 2 (F[q, p] - F[p, q]) with F[x, y] = sum_r D[x, r] h[y, r] +
sum_rst d[x, r, s, t] (y r|s t) (F[x, .] = 0 for x outside O).

Pinned in tests/test_oracle_orbopt.py by: the energy identity with the
packed cost vector and frozen-core energy of the DF intake oracle for
arbitrary primal vectors; central differences of E(kappa) for every gradient
element and second differences for every diagonal Hessian element; the
exponential against scipy.linalg.expm; and the converged v2RDM-CASSCF energy
of two-electron problems (where DQG is exact) against a brute-force CASSCF
minimisation of the full-CI energy over the non-redundant rotations.
"""
from __future__ import annotations

import numpy as np


# ---- densities ---------------------------------------------------------------

def pitzer_order(irreps):
    """User index of each Pitzer (irrep-sorted, stable) active orbital."""
    return np.argsort(np.asarray(irreps), kind="stable")


def spatial_rdms(P, x):
    """Dense spin-free D1 (n, n) and d (n, n, n, n) of the primal x, Pitzer order.

    P is an oracle.pyoracle.OracleProblem.  d[p, q, r, s] = D2ab(pr, qs) +
    D2ab(rp, sq) + s_pr s_qs (D2aa + D2bb)(pr, qs) (v2rdm_solver.cc:3110-3190
    packs the same sums, with its own factors, for the focas layout)."""
    t = P.tables()
    off = P.offsets()
    n, H = P.nmo, P.nirrep
    irr = np.asarray(P.irreps)[pitzer_order(P.irreps)]
    amopi = np.bincount(irr, minlength=H)
    poff = np.concatenate([[0], np.cumsum(amopi)[:-1]])
    gab, gaa = t["gems_ab"], t["gems_aa"]
    iab, iaa = t["ibas_ab"], t["ibas_aa"]
    D1 = np.zeros((n, n))
    for h in range(H):
        m = amopi[h]
        for i in range(m):
            for j in range(m):
                D1[poff[h] + i, poff[h] + j] = (x[off["d1a"][h] + i * m + j] + x[off["d1b"][h] + i * m + j])

    def ab(i, j, k, l):
        h = irr[i] ^ irr[j]
        if irr[k] ^ irr[l] != h:
            return 0.0
        return x[off["d2ab"][h] + iab[h, i, j] * gab[h] + iab[h, k, l]]

    def aa(i, j, k, l):
        h = irr[i] ^ irr[j]
        if i == j or k == l or irr[k] ^ irr[l] != h:
            return 0.0
        s = (-1.0 if i > j else 1.0) * (-1.0 if k > l else 1.0)
        e = iaa[h, i, j] * gaa[h] + iaa[h, k, l]
        return s * (x[off["d2aa"][h] + e] + x[off["d2bb"][h] + e])

    d = np.zeros((n, n, n, n))
    for p in range(n):
        for q in range(n):
            for r in range(n):
                for s in range(n):
                    d[p, q, r, s] = ab(p, r, q, s) + ab(r, p, s, q) + aa(p, r, q, s)
    return D1, d


def occupied_rdms(D1, d, ncore, symmetrise=True):
    """D and the fully symmetrised d over O = core + active (core first).

    Closed-shell core i, j and active t, u: D_ij = 2 delta_ij; d_ijkl =
    4 d_ij d_kl - 2 d_il d_jk; d_ijtu = d_tuij = 2 d_ij D_tu; d_ituj = d_ujit =
    d_tiju = d_juti = -d_ij D_tu; every other pattern with core indices is 0."""
    n = D1.shape[0]
    no = ncore + n
    c = np.arange(ncore)
    a = np.arange(ncore, no)
    D = np.zeros((no, no))
    D[c, c] = 2.0
    D[np.ix_(a, a)] = D1
    dd = np.zeros((no,) * 4)
    dd[np.ix_(a, a, a, a)] = d
    I = np.eye(ncore)
    dd[np.ix_(c, c, c, c)] = 4 * np.einsum("ij,kl->ijkl", I, I) - 2 * np.einsum("il,jk->ijkl", I, I)
    ID = np.einsum("ij,tu->ijtu", I, D1)
    dd[np.ix_(c, c, a, a)] = 2 * ID
    dd[np.ix_(a, a, c, c)] = 2 * ID.transpose(2, 3, 0, 1)
    dd[np.ix_(c, a, a, c)] = -ID.transpose(0, 2, 3, 1)          # d_ituj = -d_ij D_ut
    dd[np.ix_(a, c, c, a)] = -ID.transpose(2, 0, 1, 3)          # d_tiju = -d_ij D_tu
    if not symmetrise:
        return D, dd
    dsym = 0.25 * (dd + dd.transpose(1, 0, 2, 3) + dd.transpose(0, 1, 3, 2) + dd.transpose(1, 0, 3, 2))
    return D, dsym


# ---- integrals ---------------------------------------------------------------

def unpack_df(qmo, nfull):
    """(nq, ntri) lower-triangle pair order -> dense (nq, nfull, nfull)."""
    nq = qmo.shape[0]
    il = np.tril_indices(nfull)
    B = np.zeros((nq, nfull, nfull))
    pos = il[0] * (il[0] + 1) // 2 + il[1]
    B[:, il[0], il[1]] = qmo[:, pos]
    B[:, il[1], il[0]] = qmo[:, pos]
    return B


def pack_df(B):
    nfull = B.shape[1]
    il = np.tril_indices(nfull)
    pos = il[0] * (il[0] + 1) // 2 + il[1]
    out = np.zeros((B.shape[0], nfull * (nfull + 1) // 2))
    out[:, pos] = B[:, il[0], il[1]]
    return out


def rotation_pairs(nfull, core, active, full_irreps=None, active_active=False):
    """Non-redundant rotations (p, q), p-major, q ascending: virtual-core,
    virtual-active, active-core between orbitals of one irrep (the classes of
    orbopt_gradient.cc:46-146), plus active-active (p > q) on request."""
    cls = np.full(nfull, 2)
    cls[list(core)] = 0
    cls[list(active)] = 1
    irr = np.zeros(nfull, dtype=int) if full_irreps is None else np.asarray(full_irreps)
    out = []
    for p in range(nfull):
        for q in range(nfull):
            if irr[p] != irr[q] or p == q:
                continue
            if cls[p] > cls[q] or (active_active and cls[p] == cls[q] == 1 and p > q):
                out.append((p, q))
    return np.array(out, dtype=np.int64).reshape(-1, 2)


def energy(h, B, D, d, occ):
    """sum_pq D_pq h_pq + 1/2 sum d_pqrs (pq|rs) over O (no nuclear repulsion)."""
    o = np.asarray(occ)
    BO = B[:, o][:, :, o]
    g = np.einsum("Qpq,Qrs->pqrs", BO, BO)
    return float(np.sum(D * h[np.ix_(o, o)]) + 0.5 * np.sum(d * g))


def generalized_fock(h, B, D, d, occ):
    """F[x, y] (x in O, y in the full space)."""
    o = np.asarray(occ)
    BO = B[:, o][:, :, o]                                   # (Q, r, s)
    Z = np.einsum("xrst,Qst->xrQ", d, BO)
    BFO = B[:, :, o]                                        # (Q, y, r)
    return D @ h[:, o].T + np.einsum("xrQ,Qyr->xy", Z, BFO)


def gradient_hessian(h, B, D, d, occ, pairs):
    """Gradient dE/dkappa and exact diagonal d2E/dkappa2 for every pair."""
    nfull = h.shape[0]
    o = np.asarray(occ)
    pos = np.full(nfull, -1)
    pos[o] = np.arange(len(o))
    F = generalized_fock(h, B, D, d, occ)
    BO = B[:, o][:, :, o]
    gO = np.einsum("Qpq,Qrs->pqrs", BO, BO)
    Jd = np.einsum("Qp,Qtu->ptu", np.einsum("Qpp->Qp", B), BO)      # (pp|tu)
    BFO = B[:, :, o]
    Kd = np.einsum("Qps,Qpu->psu", BFO, BFO)                        # (ps|pu)
    no = len(o)
    zero2 = np.zeros((no, no))
    g = np.zeros(len(pairs))
    H = np.zeros(len(pairs))
    for k, (p, q) in enumerate(pairs):
        ip, iq = pos[p], pos[q]
        Fpq = F[ip, q] if ip >= 0 else 0.0
        Fqp = F[iq, p] if iq >= 0 else 0.0
        g[k] = 2.0 * (Fqp - Fpq)
        Fpp = F[ip, p] if ip >= 0 else 0.0
        Fqq = F[iq, q] if iq >= 0 else 0.0
        Dpp = D[ip, ip] if ip >= 0 else 0.0
        Dqq = D[iq, iq] if iq >= 0 else 0.0
        Dpq = D[ip, iq] if ip >= 0 and iq >= 0 else 0.0
        val = -2.0 * (Fpp + Fqq) + 2.0 * (Dpp * h[q, q] + Dqq * h[p, p] - 2.0 * Dpq * h[p, q])
        dpp = d[ip, ip] if ip >= 0 else zero2
        dqq = d[iq, iq] if iq >= 0 else zero2
        g12 = np.sum(dpp * Jd[q]) + np.sum(dqq * Jd[p])
        dpsp = d[ip, :, ip, :] if ip >= 0 else zero2
        dqsq = d[iq, :, iq, :] if iq >= 0 else zero2
        g13 = np.sum(dpsp * Kd[q]) + np.sum(dqsq * Kd[p])
        if ip >= 0 and iq >= 0:
            g12 -= 2.0 * np.sum(d[ip, iq] * gO[ip, iq])
            g13 -= 2.0 * np.sum(d[ip, :, iq, :] * gO[iq, :, ip, :])
        H[k] = val + 2.0 * g12 + 4.0 * g13
    return g, H


def kappa_matrix(kappa, pairs, nfull):
    K = np.zeros((nfull, nfull))
    if len(pairs):
        K[pairs[:, 0], pairs[:, 1]] = kappa
        K[pairs[:, 1], pairs[:, 0]] = -kappa
    return K


def expm(K, order=14):
    """exp(K) by scaling to ||K||_1 <= 1/2, the Taylor polynomial of degree
    14 by Horner's rule and squaring."""
    nrm = np.abs(K).sum(axis=0).max() if K.size else 0.0
    s = 0
    while nrm / (1 << s) > 0.5:
        s += 1
    P = K / float(1 << s)
    eye = np.eye(K.shape[0])
    U = eye.copy()
    for k in range(order, 0, -1):
        U = eye + (P @ U) / k
    for _ in range(s):
        U = U @ U
    return U


def rotate(h, B, U):
    return U.T @ h @ U, np.einsum("ap,Qab,bq->Qpq", U, B, U, optimize=True)


# ---- optimize_orbitals -------------------------------------------------------

def optimize_orbitals(h, B, D, d, occ, pairs, g_convergence=1e-4, e_convergence=1e-8, maxiter=10):
    """One call of OrbitalOptimizer::optimize_orbitals (orbopt_optimize.cc:25-110).

    Returns (h', B', U, info) with U the product of the applied exponentials
    (new orbitals = old orbitals U) and info = dict(iterations, gnorm, dE,
    converged, energy, trace) where trace lists (E_new, dE, gnorm, r) per
    iteration."""
    nfull = h.shape[0]
    nrot = len(pairs)
    Uacc = np.eye(nfull)

    def apply(kappa):
        nonlocal h, B, Uacc
        U = expm(kappa_matrix(kappa, pairs, nfull))
        h, B = rotate(h, B, U)
        Uacc = Uacc @ U

    eval_G = True
    converged_e = converged_g = False
    r = 1.0
    kap_cur = np.zeros(nrot)
    kap_save = np.zeros(nrot)
    kap_ref = np.zeros(nrot)
    E_init = initial_E = 0.0
    dE_quad = dE = 0.0
    gnorm = 0.0
    trace = []
    iters = 0
    for it in range(maxiter):
        if eval_G:
            gvec, Hd = gradient_hessian(h, B, D, d, occ, pairs)
            gnorm = float(np.sqrt(gvec @ gvec))
            E_init = energy(h, B, D, d, occ)
            if it == 0:
                initial_E = E_init
            # Compute_Quadratic_dE uses kappa_current before the new step
            dE_quad = 0.5 * (2.0 * gvec @ kap_cur + np.sum(kap_cur * Hd * kap_cur))
            Hc = np.abs(Hd)
            Hc[Hc < 1e-2] = 1e-2
            kap_ref = -gvec / Hc
            kap_cur = r * kap_ref
            kap_save = kap_cur.copy()
        apply(kap_cur)
        E_new = energy(h, B, D, d, occ)
        dE = E_new - E_init
        trace.append((E_new, dE, gnorm, r))
        if gnorm < g_convergence:
            converged_g = True
        if abs(dE) < e_convergence:
            converged_e = True
        iters += 1
        if converged_e or converged_g or it == maxiter - 1:
            break
        # Update_r
        eval_G = not (dE > 0.0)
        ratio = dE / dE_quad if dE_quad != 0.0 else np.inf * np.sign(dE)
        if dE < 0.0:
            if ratio > 0.75:
                r *= 1.2
            if ratio < 0.25:
                r *= 0.7
        else:
            r *= 0.7
        if not eval_G:
            kap_new = r * kap_ref - kap_cur
            kap_cur = kap_new
            kap_save = kap_save + kap_cur
    if dE > 0.0:
        if converged_e or converged_g:
            apply(-kap_cur)
        if not eval_G:
            apply(-kap_save)
    final_E = energy(h, B, D, d, occ)
    info = dict(iterations=iters, gnorm=gnorm, dE=final_E - initial_E,
                converged=bool(converged_e and converged_g), energy=final_E, trace=trace)
    return h, B, Uacc, info


def call_converged(info, g_convergence=1e-4, e_convergence=1e-8):
    """The call-level test of RotateOrbitals (v2rdm_solver.cc:3271-3274): the
    energy change of the whole optimize_orbitals call and its last gradient
    norm, not the optimiser's per-step flags."""
    return abs(info["dE"]) < e_convergence and info["gnorm"] < g_convergence


def casscf_loop(run, h, B, core, active, pairs, orbopt_frequency=500, g_convergence=1e-4, e_convergence=1e-8,
                maxiter=10, max_macro=1000):
    """The v2RDM-CASSCF loop of compute_energy (v2rdm_solver.cc:1330-1361):
    do { sdp_->solve(ORBOPT_FREQUENCY iterations); RotateOrbitals() } while
    (!orbopt_converged_ || !sdp_->is_converged()), with orbopt_converged_ set
    by the first call that passes the call-level test and never reset.

    run: an SdpRun of the active-space problem (its x is the starting point).
    Returns dict(energy = c . x + E_fzc after the last orbital call, macro =
    orbital calls, sdp_converged, orbopt_converged, h, B, oiter)."""
    from .pyoracle import df_integrals
    P = run.P
    occ = list(core) + list(active)
    oo_conv = False
    macro = 0
    sdp_conv = False
    info = None
    while macro < max_macro:
        ha, ga, efzc = df_integrals(h, pack_df(B), core, active)
        sdp_conv = run.solve(P.pack_c(ha, ga), orbopt_frequency)
        D1, d = spatial_rdms(P, run.x)
        D, dd = occupied_rdms(D1, d, len(core))
        h, B, _, info = optimize_orbitals(h, B, D, dd, occ, pairs, g_convergence, e_convergence, maxiter)
        macro += 1
        if call_converged(info, g_convergence, e_convergence):
            oo_conv = True
        if oo_conv and sdp_conv:
            break
    return dict(energy=info["energy"], macro=macro, sdp_converged=sdp_conv, orbopt_converged=oo_conv, h=h, B=B,
                oiter=run.status()["oiter"])
