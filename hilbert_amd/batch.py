"""A batch of independent DQG v2RDM SDPs solved together on one MI355X
(v2h_batch_* in include/v2rdm_hip.h)."""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _native as N


@dataclass
class Problem:
    """One fixed-orbital DQG problem; irreps are per orbital, caller order."""
    nmo: int
    nalpha: int
    nbeta: int
    irreps: object = None
    nirrep: int = None
    multiplicity: int = 0
    q2: bool = True
    g2: bool = True
    spin: bool = True
    sz: bool = True
    t1: bool = False          # POSITIVITY DQGT1 (and DQGT1T2)
    t2: bool = False          # POSITIVITY DQGT2 (and DQGT1T2)
    d3: bool = False          # CONSTRAIN_D3


def options(e_convergence=1e-6, r_convergence=1e-5, cg_convergence=1e-9, cg_maxiter=10000, maxiter=50,
            mu_update_frequency=1000, mu0=1.0) -> N.SdpOptions:
    return N.SdpOptions(e_convergence, r_convergence, cg_convergence, cg_maxiter, maxiter,
                        mu_update_frequency, mu0)


# one record of the FCIDUMP files (v2h_batch_write_fcidump)
FCIDUMP_RECORD = np.dtype([("value", "<f8"), ("labels", "<i4", (4,))])


class Batch:
    def __init__(self, problems, opts: N.SdpOptions = None, device: int = 0):
        problems = list(problems)
        descs = (N.ProblemDesc * len(problems))()
        keep = []
        for k, p in enumerate(problems):
            d, arr = N.make_desc(p.nmo, p.nalpha, p.nbeta, p.irreps, p.nirrep, p.multiplicity,
                                 p.q2, p.g2, p.spin, p.sz, getattr(p, "t1", False), getattr(p, "t2", False),
                                 getattr(p, "d3", False))
            descs[k] = d
            keep.append(arr)
        self.problems = problems
        self.opts = opts or options()
        h = C.c_void_p()
        N.check(N.lib().v2h_batch_create(descs, len(problems), C.byref(self.opts), int(device), C.byref(h)))
        self._h = h
        self._df = {}
        self.count = len(problems)
        self.n_primal, self.n_dual, self.n_blocks = [], [], []
        for i in range(self.count):
            a, b, c = C.c_int64(), C.c_int64(), C.c_int()
            N.check(N.lib().v2h_batch_sizes(self._h, i, C.byref(a), C.byref(b), C.byref(c)))
            self.n_primal.append(a.value)
            self.n_dual.append(b.value)
            self.n_blocks.append(c.value)

    def close(self):
        if getattr(self, "_h", None):
            N.lib().v2h_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # ---- setup ---------------------------------------------------------------
    def set_options(self, opts: N.SdpOptions):
        self.opts = opts
        N.check(N.lib().v2h_batch_set_options(self._h, C.byref(opts)))

    CG_TOLERANCE_FACTOR = 0.003  # the default, the oracle's

    def set_cg_tolerance_factor(self, factor: float):
        """CG stopping rule of each outer iteration: relative residual below
        factor x max(primal error, dual error), capped at 1e-2, at least
        CG_CONVERGENCE.  Changes the trajectory, not the optimum."""
        N.check(N.lib().v2h_batch_set_cg_tolerance_factor(self._h, float(factor)))

    def set_integrals(self, i, h, g):
        n = self.problems[i].nmo
        h = np.ascontiguousarray(h, dtype=np.float64).reshape(n * n)
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(n ** 4)
        N.check(N.lib().v2h_batch_set_integrals(self._h, i, h, g))

    def set_df_integrals(self, i, h_full, qmo, core, active) -> float:
        """Density-fitted Hamiltonian contracted on the device; returns E_fzc.
        The integrals stay resident for optimize_orbitals.

        h_full: (nfull, nfull); qmo: (nq, nfull (nfull + 1) / 2), the packed
        lower-triangle (Q|mn) layout of the reference's Qmo_; core: full-space
        indices of the doubly occupied core; active: full-space index of each
        active orbital, in this instance's orbital order."""
        h_full = np.ascontiguousarray(h_full, dtype=np.float64)
        nfull = h_full.shape[0]
        qmo = np.ascontiguousarray(qmo, dtype=np.float64)
        if h_full.shape != (nfull, nfull) or qmo.ndim != 2 or qmo.shape[1] != nfull * (nfull + 1) // 2:
            raise ValueError("h_full must be (nfull, nfull) and qmo (nq, nfull (nfull + 1) / 2)")
        core = np.ascontiguousarray(core, dtype=np.int32).reshape(-1)
        active = np.ascontiguousarray(active, dtype=np.int32).reshape(-1)
        if len(active) != self.problems[i].nmo:
            raise ValueError("active must list one full-space index per active orbital")
        e = np.zeros(1)
        N.check(N.lib().v2h_batch_set_df_integrals(self._h, i, nfull, h_full.reshape(-1), qmo.shape[0],
                                                   qmo.reshape(-1), len(core), core, active, e))
        self._df[i] = (nfull, qmo.shape[0])
        return float(e[0])

    def optimize_orbitals(self, i, full_irreps=None, g_convergence=1e-4, e_convergence=1e-8, maxiter=10,
                          active_active=False) -> dict:
        """One orbital-optimisation call (RotateOrbitals with the DF optimiser)
        on instance i's resident DF integrals and current primal vector.

        Rotates the integrals on the device, repacks c and E_fzc and marks the
        instance unconverged.  Returns dict(iterations, converged, nrot,
        gnorm, dE, energy = c.x + E_fzc, efzc)."""
        if i not in self._df:
            raise ValueError("instance %d has no density-fitted integrals (set_df_integrals)" % i)
        irr = None
        if full_irreps is not None:
            irr = np.ascontiguousarray(full_irreps, dtype=np.int32).reshape(-1)
            if len(irr) != self._df[i][0]:
                raise ValueError("full_irreps must have one label per orbital")
        opt = N.OrbOptOptions(float(g_convergence), float(e_convergence), int(maxiter), int(bool(active_active)))
        res = N.OrbOptResult()
        N.check(N.lib().v2h_batch_optimize_orbitals(self._h, i, None if irr is None else irr.ctypes.data_as(C.c_void_p),
                                                    C.byref(opt), C.byref(res)))
        return {k: getattr(res, k) for k, _ in N.OrbOptResult._fields_}

    def orbital_gradient(self, i, full_irreps=None, active_active=False):
        """(pairs (nrot, 2), gradient, exact diagonal Hessian) at the current
        orbitals and primal vector; nothing is rotated."""
        if i not in self._df:
            raise ValueError("instance %d has no density-fitted integrals (set_df_integrals)" % i)
        irr = None if full_irreps is None else np.ascontiguousarray(full_irreps, dtype=np.int32).reshape(-1)
        ip = None if irr is None else irr.ctypes.data_as(C.c_void_p)
        nrot = C.c_int(0)
        N.check(N.lib().v2h_batch_orbital_gradient(self._h, i, ip, int(bool(active_active)), None, None, None, 0,
                                                   C.byref(nrot)))
        m = nrot.value
        pairs = np.zeros((max(m, 1), 2), dtype=np.int32)
        g = np.zeros(max(m, 1))
        H = np.zeros(max(m, 1))
        N.check(N.lib().v2h_batch_orbital_gradient(self._h, i, ip, int(bool(active_active)),
                                                   pairs.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p),
                                                   H.ctypes.data_as(C.c_void_p), m, C.byref(nrot)))
        return pairs[:m], g[:m], H[:m]

    def df_orbitals(self, i):
        """(U, h_full, qmo) of instance i after its orbital steps: new orbital r
        = sum_s old_s U[s, r] relative to the orbitals of set_df_integrals."""
        if i not in self._df:
            raise ValueError("instance %d has no density-fitted integrals (set_df_integrals)" % i)
        nfull, nq = self._df[i]
        U = np.zeros((nfull, nfull))
        h = np.zeros((nfull, nfull))
        q = np.zeros((nq, nfull * (nfull + 1) // 2))
        N.check(N.lib().v2h_batch_get_df_orbitals(self._h, i, U.ctypes.data_as(C.c_void_p),
                                                  h.ctypes.data_as(C.c_void_p), q.ctypes.data_as(C.c_void_p)))
        return U, h, q

    def solve_casscf(self, orbopt_frequency=500, full_irreps=None, monitor=None, progress=None, **orbopt) -> list:
        """v2RDM-CASSCF: alternate orbopt_frequency BPSDP iterations with one
        optimize_orbitals call per instance until the SDP and the orbitals
        are converged (the do { solve; RotateOrbitals } while loop of
        v2rdm_solver.cc:1296-1361; ORBOPT_FREQUENCY default 500).

        full_irreps: None, one array for every instance, or a list of
        per-instance arrays.  progress(macro, results) is called after every
        round of orbital calls.  Returns per instance dict(energy = c.x + E_fzc
        after the last orbital step, sdp_converged, orbopt_converged,
        macro = orbital calls, last = the last call's result)."""
        n = self.count
        if full_irreps is None or (len(full_irreps) and np.ndim(full_irreps[0]) == 0):
            irr = [full_irreps] * n
        else:
            irr = list(full_irreps)
        oo_conv = [False] * n
        sdp_conv = [False] * n
        done = [False] * n
        last = [None] * n
        macro = [0] * n
        while not all(done):
            self.solve(orbopt_frequency, monitor)
            for i in range(n):
                if done[i]:
                    continue
                st = self.status(i)
                if st["status_bits"] or st["oiter"] >= self.opts.maxiter:
                    done[i] = True
                    continue
                last[i] = self.optimize_orbitals(i, irr[i], **orbopt)
                macro[i] += 1
                # orbopt_converged_ is set by the first call that passes the
                # call-level test and never reset (v2rdm_solver.cc:3271-3274);
                # the loop ends once it is set and the SDP chunk before the
                # call converged (the while test of v2rdm_solver.cc:1361)
                if last[i]["converged"]:
                    oo_conv[i] = True
                if oo_conv[i] and st["converged"]:
                    done[i] = sdp_conv[i] = True
            if progress:
                progress(max(macro), last)
        out = []
        for i in range(n):
            st = self.status(i)
            out.append(dict(energy=None if last[i] is None else last[i]["energy"],
                            sdp_converged=sdp_conv[i], orbopt_converged=oo_conv[i],
                            macro=macro[i], oiter=st["oiter"], last=last[i]))
        return out

    def write_fcidump(self, i, int_path, rdm_path, h_full, qmo, core, active, enuc=0.0, full_irreps=None):
        """hilbert's FCIDUMP files (integrals, RDMs of the current x) for
        instance i from DF integrals; returns the two record counts.  Each
        record: float64 value, four int32 labels (numpy dtype
        ``FCIDUMP_RECORD``)."""
        h_full = np.ascontiguousarray(h_full, dtype=np.float64)
        nfull = h_full.shape[0]
        qmo = np.ascontiguousarray(qmo, dtype=np.float64)
        if h_full.shape != (nfull, nfull) or qmo.ndim != 2 or qmo.shape[1] != nfull * (nfull + 1) // 2:
            raise ValueError("h_full must be (nfull, nfull) and qmo (nq, nfull (nfull + 1) / 2)")
        core = np.ascontiguousarray(core, dtype=np.int32).reshape(-1)
        active = np.ascontiguousarray(active, dtype=np.int32).reshape(-1)
        if len(active) != self.problems[i].nmo:
            raise ValueError("active must list one full-space index per active orbital")
        irr = None
        if full_irreps is not None:
            irr = np.ascontiguousarray(full_irreps, dtype=np.int32).reshape(-1)
            if len(irr) != nfull:
                raise ValueError("full_irreps must have one label per orbital")
        n_int, n_rdm = C.c_int64(), C.c_int64()
        N.check(N.lib().v2h_batch_write_fcidump(
            self._h, i, nfull, None if irr is None else irr.ctypes.data_as(C.c_void_p), h_full.reshape(-1),
            qmo.shape[0], qmo.reshape(-1), len(core), core, active, float(enuc), os.fsencode(int_path),
            os.fsencode(rdm_path), C.byref(n_int), C.byref(n_rdm)))
        return n_int.value, n_rdm.value

    def set_cost(self, i, c):
        N.check(N.lib().v2h_batch_set_cost(self._h, i, np.ascontiguousarray(c, dtype=np.float64)))

    def set_guess(self, i, x=None):
        if x is None:
            N.check(N.lib().v2h_batch_set_guess(self._h, i, None))
        else:
            x = np.ascontiguousarray(x, dtype=np.float64)
            N.check(N.lib().v2h_batch_set_guess(self._h, i, x.ctypes.data_as(C.c_void_p)))

    def set_guess_hf(self, i):
        """hilbert's TPDM_GUESS HF (see v2h_batch_set_guess_hf)."""
        N.check(N.lib().v2h_batch_set_guess_hf(self._h, i))

    def reset(self, i):
        N.check(N.lib().v2h_batch_reset(self._h, i))

    def block_dims(self, i):
        d = np.zeros(self.n_blocks[i], dtype=np.int32)
        N.check(N.lib().v2h_batch_block_dims(self._h, i, d))
        return d

    def orbital_order(self, i):
        o = np.zeros(self.problems[i].nmo, dtype=np.int32)
        N.check(N.lib().v2h_batch_orbital_order(self._h, i, o))
        return o

    # ---- solve ---------------------------------------------------------------
    def solve(self, maxiter, monitor=None) -> int:
        cb = N.MONITOR(lambda *a: monitor(*a[:-1])) if monitor else None
        nc = C.c_int(0)
        N.check(N.lib().v2h_batch_solve(self._h, int(maxiter), C.cast(cb, C.c_void_p) if cb else None, None,
                                        C.byref(nc)))
        return nc.value

    def solve_to_convergence(self, chunk=None, monitor=None) -> int:
        """Run until every instance converged or hit MAXITER (the
        do { solve } while (!converged) loop of v2rdm_solver.cc:1330-1361)."""
        chunk = chunk or max(1, self.opts.mu_update_frequency)
        while True:
            self.solve(chunk, monitor)
            st = [self.status(i) for i in range(self.count)]
            if all(s["converged"] or s["oiter"] >= self.opts.maxiter or s["status_bits"] for s in st):
                return sum(s["converged"] for s in st)

    def last_solve_seconds(self) -> float:
        return N.lib().v2h_batch_last_solve_seconds(self._h)

    KERNEL_KINDS = ("pre", "cg_at", "cg_a", "post", "project_jacobi16", "project_tri64", "project_tri128",
                    "project_tri192", "project_tri224", "project_tri256", "project_jacobi_big", "df_intake",
                    "pair_gram", "orbopt", "project_refine", "project_seed")

    PROJECT_TRIDIAGONAL, PROJECT_REFINED = 0, 1

    def set_projection(self, mode):
        """PROJECT_REFINED (default): blocks of 65..256 rows refine the
        previous iteration's eigenvectors on the FP64 matrix cores and fall
        back to the tridiagonal eigensolver; PROJECT_TRIDIAGONAL: always the
        tridiagonal eigensolver."""
        N.check(N.lib().v2h_batch_set_projection(self._h, int(mode)))
        self._projection = int(mode)

    def projection_mode(self) -> int:
        return getattr(self, "_projection", self.PROJECT_REFINED)

    def projection_stats(self) -> dict:
        st = np.zeros(6, dtype=np.int64)
        N.check(N.lib().v2h_batch_projection_stats(self._h, st))
        return dict(zip(("refined", "fallback", "rejected_divergence", "rejected_correction",
                         "rejected_clusters", "flops"), (int(v) for v in st)))

    def set_refine_params(self, cluster_rel=1e-8, accept_e=2e-7, accept_cpl=1e-13, max_steps=5):
        N.check(N.lib().v2h_batch_set_refine_params(self._h, float(cluster_rel), float(accept_e),
                                                    float(accept_cpl), int(max_steps)))

    def set_profiling(self, on=True):
        N.check(N.lib().v2h_batch_set_profiling(self._h, int(on)))

    def tuning(self) -> dict:
        """The launch configuration (v2h_tuning) as a dict."""
        t = N.Tuning()
        N.check(N.lib().v2h_batch_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, cfg: dict):
        """Set the launch configuration (changes no result bit)."""
        t = N.Tuning.from_dict(cfg)
        N.check(N.lib().v2h_batch_set_tuning(self._h, C.byref(t)))

    GRAPHS_OFF, GRAPHS_CG, GRAPHS_ALL = 0, 1, 2

    def set_graphs(self, mode=GRAPHS_CG):
        """hipGraph replay: GRAPHS_CG the first CG chunk of every outer
        iteration; GRAPHS_ALL also k_post and the whole projection;
        GRAPHS_OFF (default) the same launches one by one.  All three give
        the same bits."""
        N.check(N.lib().v2h_batch_set_graphs(self._h, int(mode)))

    def kernel_times(self, full=False) -> dict:
        """{kind: (total ms, launches)}; full=True: only launches that
        processed the whole batch (every instance running)."""
        n = len(self.KERNEL_KINDS)
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        fms = np.zeros(n)
        fcnt = np.zeros(n, dtype=np.int64)
        N.check(N.lib().v2h_batch_kernel_times(self._h, ms, cnt, fms, fcnt))
        if full:
            ms, cnt = fms, fcnt
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.KERNEL_KINDS)}

    def reset_kernel_times(self):
        N.check(N.lib().v2h_batch_reset_kernel_times(self._h))

    def launch_stats(self) -> dict:
        st = np.zeros(4, dtype=np.int64)
        N.check(N.lib().v2h_batch_launch_stats(self._h, st))
        return dict(zip(("launches", "a_workgroups", "at_workgroups", "blocks"), st.tolist()))

    def status(self, i) -> dict:
        s = N.Status()
        N.check(N.lib().v2h_batch_status(self._h, i, C.byref(s)))
        return {"mu": s.mu, "energy_primal": s.energy_primal, "energy_dual": s.energy_dual,
                "eps_p": s.primal_error, "eps_d": s.dual_error, "oiter": s.oiter_total,
                "iiter": s.iiter_total, "converged": bool(s.converged), "active": bool(s.active),
                "status_bits": s.status_bits, "cg_last": s.last_cg_iterations}

    # ---- vectors ---------------------------------------------------------------
    def _vec(self, fn, i, n):
        out = np.zeros(n)
        N.check(fn(self._h, i, out))
        return out

    def x(self, i):
        return self._vec(N.lib().v2h_batch_get_x, i, self.n_primal[i])

    def y(self, i):
        return self._vec(N.lib().v2h_batch_get_y, i, self.n_dual[i])

    def z(self, i):
        return self._vec(N.lib().v2h_batch_get_z, i, self.n_primal[i])

    def c(self, i):
        return self._vec(N.lib().v2h_batch_get_c, i, self.n_primal[i])

    def b(self, i):
        return self._vec(N.lib().v2h_batch_get_b, i, self.n_dual[i])

    def set_y(self, i, y):
        N.check(N.lib().v2h_batch_set_y(self._h, i, np.ascontiguousarray(y, dtype=np.float64)))

    def set_z(self, i, z):
        N.check(N.lib().v2h_batch_set_z(self._h, i, np.ascontiguousarray(z, dtype=np.float64)))

    def mu(self, i):
        m = C.c_double()
        N.check(N.lib().v2h_batch_get_mu(self._h, i, C.byref(m)))
        return m.value

    def set_mu(self, i, mu):
        N.check(N.lib().v2h_batch_set_mu(self._h, i, float(mu)))

    # ---- reduced density matrices (orbital indices in the caller's order) ------
    def opdm(self, i, type="SUM"):
        """Dense n x n 1-RDM of instance i: "A", "B" or "SUM"."""
        n = self.problems[i].nmo
        out = np.zeros((n, n))
        N.check(N.lib().v2h_batch_get_opdm(self._h, i, type.encode(), out.reshape(-1)))
        return out

    def tpdm(self, i):
        """Dense spin-free 2-RDM, shape (n, n, n, n), element [i, k, j, l] =
        D(ij, kl) as in hilbert's get_tpdm matrix [i n + k][j n + l]."""
        n = self.problems[i].nmo
        out = np.zeros(n ** 4)
        N.check(N.lib().v2h_batch_get_tpdm(self._h, i, out))
        return out.reshape(n, n, n, n)

    def _sparse(self, fn, i, type, width):
        cnt = C.c_int64()
        N.check(fn(self._h, i, type.encode(), None, None, 0, C.byref(cnt)))
        idx = np.zeros((cnt.value, width), dtype=np.int32)
        val = np.zeros(cnt.value)
        N.check(fn(self._h, i, type.encode(), idx.ctypes.data_as(C.c_void_p), val.ctypes.data_as(C.c_void_p),
                   cnt.value, C.byref(cnt)))
        return idx, val

    def opdm_sparse(self, i, type):
        """(indices [m, 2], values [m]) in the order of hilbert's get_opdm_sparse."""
        return self._sparse(N.lib().v2h_batch_get_opdm_sparse, i, type, 2)

    def tpdm_sparse(self, i, type):
        """(indices [m, 4], values [m]) in the order of hilbert's get_tpdm_sparse."""
        return self._sparse(N.lib().v2h_batch_get_tpdm_sparse, i, type, 4)

    def write_rdm_files(self, i, prefix, active, core=(), which=N.RDM_ALL_FULL):
        """Instance i's RDM files for downstream codes (hilbert's
        WriteActiveTPDM, WriteTPDM, WriteTPDMSpinFree, WriteOPDM and
        WriteTPDM_IWL; write_tpdm.cc:40-595, write_opdm.cc:40,
        write_tpdm_iwl.cc:45): <prefix>.d2aa .d2bb .d2ab .d2 .d1a .d1b (keyed
        records) and <prefix>.aa.iwl .bb.iwl .ab.iwl (IWL buffers), chosen by
        `which` (N.RDM_* flags).  `active[u]` is the full-space label of
        active orbital u, `core` the labels of the doubly occupied core
        orbitals.  Returns the record count of each file written."""
        if len(active) != self.problems[i].nmo:
            raise ValueError("one full-space label per active orbital needed")
        return N.write_rdm_files(N.lib().v2h_batch_write_rdm_files, (self._h, i), prefix, active, core, which)

    def Au(self, i, u):
        out = np.zeros(self.n_dual[i])
        N.check(N.lib().v2h_batch_Au(self._h, i, np.ascontiguousarray(u, dtype=np.float64), out))
        return out

    def ATu(self, i, y):
        out = np.zeros(self.n_primal[i])
        N.check(N.lib().v2h_batch_ATu(self._h, i, np.ascontiguousarray(y, dtype=np.float64), out))
        return out

    def cost_adjoint(self, i, xbar=None):
        n = self.problems[i].nmo
        dh = np.zeros(n * n)
        dg = np.zeros(n ** 4)
        ptr = None
        if xbar is not None:
            xbar = np.ascontiguousarray(xbar, dtype=np.float64)
            ptr = xbar.ctypes.data_as(C.c_void_p)
        N.check(N.lib().v2h_batch_cost_adjoint(self._h, i, ptr, dh, dg))
        return dh.reshape(n, n), dg.reshape(n, n, n, n)
