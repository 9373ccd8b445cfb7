"""ctypes wrapper of the CPU oracle (oracle/libv2rdm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product (hilbert_amd)
never imports it.
"""
from __future__ import annotations

import ctypes as C
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libv2rdm_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

MONITOR = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                      C.c_double, C.c_double, C.c_double, C.c_void_p)


class SdpOptions(C.Structure):
    """Mirror of orc_sdp_options (names follow hilbert's option registry)."""
    _fields_ = [("e_convergence", C.c_double), ("r_convergence", C.c_double),
                ("cg_convergence", C.c_double), ("cg_maxiter", C.c_int),
                ("maxiter", C.c_int), ("mu_update_frequency", C.c_int),
                ("mu0", C.c_double), ("nthreads", C.c_int)]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None
_dsyevd = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = C.CDLL(_LIBPATH)
        L.orc_problem_new.restype = C.c_void_p
        L.orc_problem_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_problem_delete.argtypes = [C.c_void_p]
        for f in ("orc_n_primal", "orc_n_dual"):
            getattr(L, f).restype = C.c_long
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_nblocks.restype = C.c_int
        L.orc_nblocks.argtypes = [C.c_void_p]
        L.orc_get_dims.argtypes = [C.c_void_p, _ip]
        L.orc_get_gems.argtypes = [C.c_void_p, _ip, _ip]
        L.orc_get_tables.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip]
        L.orc_get_offsets.argtypes = [C.c_void_p, _lp]
        L.orc_build_b.argtypes = [C.c_void_p, _dp]
        L.orc_pack_c.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_pack_c_adjoint.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_hubbard_integrals.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_df_integrals.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, _dp, _dp, _dp]
        L.orc_problem_enable_t.restype = C.c_int
        L.orc_problem_enable_t.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_get_t_offsets.argtypes = [C.c_void_p, _lp]
        L.orc_get_trip_counts.argtypes = [C.c_void_p, _ip, _ip, _ip]
        L.orc_fcidump_map.argtypes = [C.c_int, C.c_int, C.c_void_p, _ip, _ip]
        L.orc_fcidump_integrals.restype = C.c_long
        L.orc_fcidump_integrals.argtypes = [C.c_int, C.c_void_p, _ip, _dp, C.c_int, _dp, C.c_double, C.c_void_p,
                                            C.c_void_p]
        L.orc_fcidump_rdms.restype = C.c_long
        L.orc_fcidump_rdms.argtypes = [C.c_void_p, _dp, _ip, _ip, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_Au.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_ATu.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_guess_random.argtypes = [C.c_long, _dp]
        L.orc_spin_squared.restype = C.c_double
        L.orc_spin_squared.argtypes = [C.c_void_p, _dp]
        L.orc_opdm.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_tpdm.restype = C.c_int
        L.orc_tpdm.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_sdp_state_new.restype = C.c_void_p
        L.orc_sdp_state_new.argtypes = [C.c_long, C.c_long]
        L.orc_sdp_state_delete.argtypes = [C.c_void_p]
        L.orc_sdp_state_scalars.argtypes = [C.c_void_p, _dp, _lp]
        L.orc_sdp_state_y.restype = C.POINTER(C.c_double)
        L.orc_sdp_state_y.argtypes = [C.c_void_p]
        L.orc_sdp_state_z.restype = C.POINTER(C.c_double)
        L.orc_sdp_state_z.argtypes = [C.c_void_p]
        L.orc_solve_dqg.restype = C.c_int
        L.orc_solve_dqg.argtypes = [C.c_void_p, C.POINTER(SdpOptions), C.c_void_p, _dp, _dp, _dp,
                                    C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_project_block.restype = C.c_int
        L.orc_project_block.argtypes = [C.c_int, _dp, C.c_double, _dp, _dp, C.c_void_p, _dp,
                                        C.c_int, _ip, C.c_int]
        L.orc_jacobi_eig.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_cg_tolerance.restype = C.c_double
        L.orc_cg_tolerance.argtypes = [C.POINTER(SdpOptions), C.c_long, C.c_double, C.c_double]
        _lib = L
    return _lib


def dsyevd_pointer():
    """Address of LAPACK dsyevd from scipy's bundled OpenBLAS (LP64), or None.

    The OpenBLAS thread count is pinned to 1: the oracle parallelises over
    blocks with OpenMP instead."""
    global _dsyevd
    if _dsyevd is None:
        try:
            import scipy
            d = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
            cands = sorted(glob.glob(os.path.join(d, "libscipy_openblas*.so")))
            ob = C.CDLL(cands[0])
            try:
                ob.scipy_openblas_set_num_threads(1)
            except AttributeError:
                pass
            _dsyevd = (ob, C.cast(ob.scipy_dsyevd_, C.c_void_p).value)
        except Exception:  # pragma: no cover - scipy is part of the image
            _dsyevd = (None, None)
    return _dsyevd[1]


def df_integrals(h_full, qmo, core, active):
    """Active-space (h with core Fock terms, (pq|rs), E_fzc) from DF integrals.

    h_full: (nfull, nfull); qmo: (nq, nfull (nfull + 1) / 2) in the packed
    lower-triangle pair order; core / active: full-space orbital indices.
    Returns (h_act, g_act, efzc) in the order of ``active``."""
    h_full = np.ascontiguousarray(h_full, dtype=np.float64)
    nfull = h_full.shape[0]
    qmo = np.ascontiguousarray(qmo, dtype=np.float64)
    nq = qmo.shape[0]
    assert qmo.shape == (nq, nfull * (nfull + 1) // 2)
    core = np.ascontiguousarray(core, dtype=np.int32)
    active = np.ascontiguousarray(active, dtype=np.int32)
    n = len(active)
    h_act = np.zeros(n * n)
    g_act = np.zeros(n ** 4)
    e = np.zeros(1)
    lib().orc_df_integrals(nfull, h_full.reshape(-1), nq, qmo.reshape(-1), len(core), core, n, active, h_act,
                           g_act, e)
    return h_act.reshape(n, n), g_act.reshape(n, n, n, n), float(e[0])


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def fcidump_map(space, full_irreps=None, nirrep=1):
    """1-based FCIDUMP label of every orbital (core, active, virtual; then irrep)."""
    space = np.ascontiguousarray(space, dtype=np.int32)
    irr = None if full_irreps is None else np.ascontiguousarray(full_irreps, dtype=np.int32)
    out = np.zeros(len(space), dtype=np.int32)
    lib().orc_fcidump_map(len(space), nirrep, _ptr(irr), space, out)
    return out


def fcidump_integrals(h_full, qmo, enuc, lab_map, full_irreps=None):
    """(values, labels[:, 4]) of the integral file, in writing order."""
    h_full = np.ascontiguousarray(h_full, dtype=np.float64)
    nfull = h_full.shape[0]
    qmo = np.ascontiguousarray(qmo, dtype=np.float64)
    irr = None if full_irreps is None else np.ascontiguousarray(full_irreps, dtype=np.int32)
    lab_map = np.ascontiguousarray(lab_map, dtype=np.int32)
    n = lib().orc_fcidump_integrals(nfull, _ptr(irr), lab_map, h_full.reshape(-1), qmo.shape[0], qmo.reshape(-1),
                                    float(enuc), None, None)
    val = np.zeros(n)
    lab = np.zeros((n, 4), dtype=np.int32)
    lib().orc_fcidump_integrals(nfull, _ptr(irr), lab_map, h_full.reshape(-1), qmo.shape[0], qmo.reshape(-1),
                                float(enuc), _ptr(val), _ptr(lab))
    return val, lab


class OracleProblem:
    """One DQG problem in the oracle (tables, b, c-packing, maps, solve)."""

    def __init__(self, nmo, nalpha, nbeta, irreps=None, nirrep=None, multiplicity=None,
                 q2=True, g2=True, spin=True, sz=True, t1=False, t2=False, d3=False):
        L = lib()
        if irreps is None:
            irreps = [0] * nmo
        irreps = np.ascontiguousarray(irreps, dtype=np.int32)
        if nirrep is None:
            nirrep = 1
            while nirrep <= int(irreps.max()):
                nirrep *= 2
        if multiplicity is None:
            multiplicity = nalpha - nbeta + 1
        self.nmo, self.nalpha, self.nbeta = nmo, nalpha, nbeta
        self.nirrep, self.irreps, self.multiplicity = nirrep, irreps, multiplicity
        self.flags = (q2, g2, spin, sz)
        self._p = L.orc_problem_new(nmo, nalpha, nbeta, multiplicity, irreps, nirrep,
                                    int(q2), int(g2), int(spin), int(sz))
        if not self._p:
            raise ValueError("invalid oracle problem")
        self.t1, self.t2, self.d3 = bool(t1), bool(t2), bool(d3)
        if (t1 or t2 or d3) and L.orc_problem_enable_t(self._p, int(t1), int(t2), int(d3)) != 0:
            raise ValueError("T1 and T2 need Q2 and G2")
        self.n_primal = L.orc_n_primal(self._p)
        self.n_dual = L.orc_n_dual(self._p)
        nb = L.orc_nblocks(self._p)
        self.dims = np.zeros(nb, dtype=np.int32)
        L.orc_get_dims(self._p, self.dims)

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_problem_delete(self._p)
            self._p = None

    # ---- tables -------------------------------------------------------------
    def gems(self):
        gab = np.zeros(self.nirrep, dtype=np.int32)
        gaa = np.zeros(self.nirrep, dtype=np.int32)
        lib().orc_get_gems(self._p, gab, gaa)
        return gab, gaa

    def tables(self):
        gab, gaa = self.gems()
        n, H = self.nmo, self.nirrep
        bab = np.zeros(2 * int(gab.sum()) + 1, dtype=np.int32)
        baa = np.zeros(2 * int(gaa.sum()) + 1, dtype=np.int32)
        iab = np.zeros(H * n * n, dtype=np.int32)
        iaa = np.zeros(H * n * n, dtype=np.int32)
        lib().orc_get_tables(self._p, bab, iab, baa, iaa)
        return {"gems_ab": gab, "gems_aa": gaa,
                "bas_ab": bab[:2 * int(gab.sum())].reshape(-1, 2),
                "bas_aa": baa[:2 * int(gaa.sum())].reshape(-1, 2),
                "ibas_ab": iab.reshape(H, n, n), "ibas_aa": iaa.reshape(H, n, n)}

    OFFSET_NAMES = ("d2ab", "d2aa", "d2bb", "d200", "d1a", "d1b", "q1a", "q1b",
                    "q2ab", "q2aa", "q2bb", "g2ab", "g2ba", "g2aa")

    def offsets(self):
        off = np.zeros(14 * self.nirrep, dtype=np.int64)
        lib().orc_get_offsets(self._p, off)
        off = off.reshape(14, self.nirrep)
        return {k: off[i].copy() for i, k in enumerate(self.OFFSET_NAMES)}

    # ---- vectors ------------------------------------------------------------
    def b(self):
        out = np.zeros(self.n_dual)
        lib().orc_build_b(self._p, out)
        return out

    def pack_c(self, h, g):
        n = self.nmo
        h = np.ascontiguousarray(h, dtype=np.float64).reshape(n * n)
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(n ** 4)
        out = np.zeros(self.n_primal)
        lib().orc_pack_c(self._p, h, g, out)
        return out

    def t1_offsets(self):
        """{"t1aaa", "t1bbb", "t1aab", "t1bba", "t2aaa", "t2bbb", "t2aab",
        "t2bba"}: per-irrep block offsets."""
        H = self.nirrep
        off = np.zeros(12 * H, dtype=np.int64)
        lib().orc_get_t_offsets(self._p, off)
        names = ("t1aaa", "t1bbb", "t1aab", "t1bba", "t2aaa", "t2bbb", "t2aab", "t2bba",
                 "d3aaa", "d3bbb", "d3aab", "d3bba")
        return {k: off[a * H:(a + 1) * H] for a, k in enumerate(names)}

    def trip_counts(self, aba=False):
        taab = np.zeros(self.nirrep, dtype=np.int32)
        taaa = np.zeros(self.nirrep, dtype=np.int32)
        taba = np.zeros(self.nirrep, dtype=np.int32)
        lib().orc_get_trip_counts(self._p, taab, taaa, taba)
        return (taab, taaa, taba) if aba else (taab, taaa)

    def fcidump_rdms(self, x, act_full, lab_map, ninact):
        """(values, labels[:, 4]) of the RDM file, in writing order."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        act_full = np.ascontiguousarray(act_full, dtype=np.int32)
        lab_map = np.ascontiguousarray(lab_map, dtype=np.int32)
        n = lib().orc_fcidump_rdms(self._p, x, act_full, lab_map, int(ninact), None, None)
        val = np.zeros(n)
        lab = np.zeros((n, 4), dtype=np.int32)
        lib().orc_fcidump_rdms(self._p, x, act_full, lab_map, int(ninact), _ptr(val), _ptr(lab))
        return val, lab

    def pack_c_adjoint(self, xbar):
        n = self.nmo
        hb = np.zeros(n * n)
        gb = np.zeros(n ** 4)
        lib().orc_pack_c_adjoint(self._p, np.ascontiguousarray(xbar, dtype=np.float64), hb, gb)
        return hb.reshape(n, n), gb.reshape(n, n, n, n)

    def Au(self, u):
        out = np.zeros(self.n_dual)
        lib().orc_Au(self._p, np.ascontiguousarray(u, dtype=np.float64), out)
        return out

    def ATu(self, y):
        out = np.zeros(self.n_primal)
        lib().orc_ATu(self._p, np.ascontiguousarray(y, dtype=np.float64), out)
        return out

    def guess(self):
        out = np.zeros(self.n_primal)
        lib().orc_guess_random(self.n_primal, out)
        return out

    def spin_squared(self, x):
        return lib().orc_spin_squared(self._p, np.ascontiguousarray(x, dtype=np.float64))

    def opdm(self, x):
        out = np.zeros(self.nmo * self.nmo)
        lib().orc_opdm(self._p, np.ascontiguousarray(x, dtype=np.float64), out)
        return out.reshape(self.nmo, self.nmo)

    def tpdm(self, x):
        n = self.nmo
        out = np.zeros(n ** 4)
        if lib().orc_tpdm(self._p, np.ascontiguousarray(x, dtype=np.float64), out) != 0:
            raise ValueError("tpdm export is C1 only")
        return out.reshape(n * n, n * n)

    def run(self, x=None, **opts):
        """A BPSDP run whose y, z and mu persist across chunks (see SdpRun)."""
        return SdpRun(self, x, **opts)

    # ---- solve --------------------------------------------------------------
    def solve(self, c, x=None, e_convergence=1e-6, r_convergence=1e-5, cg_convergence=1e-9,
              cg_maxiter=10000, maxiter=20000, mu_update_frequency=1000, mu0=1.0,
              nthreads=None, use_lapack=True, chunk=None, monitor=None):
        """Run BPSDP to convergence (or maxiter).  Returns a dict."""
        L = lib()
        if x is None:
            x = self.guess()
        x = np.array(x, dtype=np.float64, copy=True)
        b = self.b()
        c = np.ascontiguousarray(c, dtype=np.float64)
        o = SdpOptions(e_convergence, r_convergence, cg_convergence, cg_maxiter, maxiter,
                       mu_update_frequency, mu0, nthreads or (os.cpu_count() or 1))
        eig = dsyevd_pointer() if use_lapack else None
        mon = MONITOR(monitor) if monitor else None
        s = L.orc_sdp_state_new(self.n_primal, self.n_dual)
        try:
            chunk = chunk or maxiter
            conv = 0
            while True:
                conv = L.orc_solve_dqg(s, C.byref(o), self._p, x, b, c, chunk,
                                       eig, C.cast(mon, C.c_void_p) if mon else None, 1)
                sc = np.zeros(7)
                cnt = np.zeros(3, dtype=np.int64)
                L.orc_sdp_state_scalars(s, sc, cnt)
                if conv or cnt[0] >= maxiter:
                    break
            y = np.ctypeslib.as_array(L.orc_sdp_state_y(s), shape=(self.n_dual,)).copy()
            z = np.ctypeslib.as_array(L.orc_sdp_state_z(s), shape=(self.n_primal,)).copy()
        finally:
            L.orc_sdp_state_delete(s)
        return {"x": x, "y": y, "z": z, "converged": bool(conv), "mu": sc[0],
                "energy_primal": sc[1], "energy_dual": sc[2], "eps_p": sc[3], "eps_d": sc[4],
                "cg_time": sc[5], "proj_time": sc[6], "oiter": int(cnt[0]), "iiter": int(cnt[1])}


class SdpRun:
    """BPSDP whose y, z and mu persist across solve() chunks, as libsdp keeps
    them between sdp_->solve calls (checkpoint.cc:58-120), for loops that
    change c between chunks (the v2RDM-CASSCF loop, v2rdm_solver.cc:1330-1361).
    Every chunk iterates at least once and re-derives the norms of b and c."""

    def __init__(self, P, x=None, e_convergence=1e-6, r_convergence=1e-5, cg_convergence=1e-9,
                 cg_maxiter=10000, maxiter=20000, mu_update_frequency=1000, mu0=1.0, nthreads=None):
        self.P = P
        self.x = np.array(P.guess() if x is None else x, dtype=np.float64, copy=True)
        self.b = P.b()
        self.o = SdpOptions(e_convergence, r_convergence, cg_convergence, cg_maxiter, maxiter,
                            mu_update_frequency, mu0, nthreads or (os.cpu_count() or 1))
        self.s = lib().orc_sdp_state_new(P.n_primal, P.n_dual)
        self.converged = False

    def solve(self, c, maxiter):
        c = np.ascontiguousarray(c, dtype=np.float64)
        self.converged = bool(lib().orc_solve_dqg(self.s, C.byref(self.o), self.P._p, self.x, self.b, c,
                                                  int(maxiter), dsyevd_pointer(), None, 1))
        return self.converged

    def status(self):
        sc = np.zeros(7)
        cnt = np.zeros(3, dtype=np.int64)
        lib().orc_sdp_state_scalars(self.s, sc, cnt)
        return {"mu": sc[0], "energy_primal": sc[1], "energy_dual": sc[2], "eps_p": sc[3], "eps_d": sc[4],
                "oiter": int(cnt[0]), "iiter": int(cnt[1]), "converged": self.converged}

    def close(self):
        if getattr(self, "s", None):
            lib().orc_sdp_state_delete(self.s)
            self.s = None

    def __del__(self):
        self.close()
