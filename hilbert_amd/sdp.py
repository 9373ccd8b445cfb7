"""BPSDP with caller-supplied constraint maps: the generic libsdp contract
(BPSDPSolver(n_primal, n_dual, options); solve(x, b, c, block_dims, maxiter,
Au, ATu, monitor), v2rdm_solver.cc:433, 1340) over v2h_sdp_* in
include/v2rdm_hip.h.  The maps are Python callables on numpy vectors; the
PSD projection runs on the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

MAP = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


class BPSDPSolver:
    def __init__(self, n_primal, n_dual, opts: N.SdpOptions = None, device: int = 0):
        from .batch import options
        self.n_primal, self.n_dual = int(n_primal), int(n_dual)
        self.opts = opts or options()
        h = C.c_void_p()
        N.check(N.lib().v2h_sdp_create(self.n_primal, self.n_dual, C.byref(self.opts), int(device), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            N.lib().v2h_sdp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def solve(self, x, b, c, block_dims, maxiter, Au, ATu, monitor=None, print_level=1, data=None) -> bool:
        """Au(u) -> A u (n_dual), ATu(y) -> A^T y (n_primal); x is updated in place.

        monitor(print_level, oiter, iiter, E_p, E_d, mu, eps_p, eps_d, data)
        is libsdp's progress monitor (v2rdm_solver.cc:115-124), called after
        every outer iteration with print_level and data passed through, as
        sdp_->solve(..., bpsdp_monitor, 1, this) does (v2rdm_solver.cc:1340)."""
        npr, nd = self.n_primal, self.n_dual

        def wrap(fn, n_in, n_out):
            def cb(out, inp, _):
                o = np.ctypeslib.as_array(out, shape=(n_out,))
                i = np.ctypeslib.as_array(inp, shape=(n_in,))
                o[:] = fn(i)
            return MAP(cb)

        cau, catu = wrap(Au, npr, nd), wrap(ATu, nd, npr)
        mon = N.SDP_MONITOR(lambda *a: monitor(*a[:-1], data)) if monitor else None
        conv = C.c_int(0)
        N.check(N.lib().v2h_sdp_solve(self._h, x, np.ascontiguousarray(b, dtype=np.float64),
                                      np.ascontiguousarray(c, dtype=np.float64),
                                      np.ascontiguousarray(block_dims, dtype=np.int32), len(block_dims),
                                      int(maxiter), C.cast(cau, C.c_void_p), C.cast(catu, C.c_void_p),
                                      C.cast(mon, C.c_void_p) if mon else None, int(print_level), None,
                                      C.byref(conv)))
        return bool(conv.value)

    def status(self) -> dict:
        s = N.Status()
        N.check(N.lib().v2h_sdp_status(self._h, C.byref(s)))
        return {"mu": s.mu, "energy_primal": s.energy_primal, "energy_dual": s.energy_dual,
                "eps_p": s.primal_error, "eps_d": s.dual_error, "oiter": s.oiter_total,
                "iiter": s.iiter_total, "converged": bool(s.converged), "status_bits": s.status_bits}

    def is_converged(self) -> bool:
        return self.status()["converged"]

    def iiter_total(self) -> int:
        return self.status()["iiter"]

    def oiter_total(self) -> int:
        return self.status()["oiter"]

    def get_y(self):
        y = np.zeros(self.n_dual)
        N.check(N.lib().v2h_sdp_get_y(self._h, y))
        return y

    def get_z(self):
        z = np.zeros(self.n_primal)
        N.check(N.lib().v2h_sdp_get_z(self._h, z))
        return z

    def set_y(self, y):
        N.check(N.lib().v2h_sdp_set_y(self._h, np.ascontiguousarray(y, dtype=np.float64)))

    def set_z(self, z):
        N.check(N.lib().v2h_sdp_set_z(self._h, np.ascontiguousarray(z, dtype=np.float64)))

    def get_mu(self) -> float:
        m = C.c_double()
        N.check(N.lib().v2h_sdp_get_mu(self._h, C.byref(m)))
        return m.value

    def set_mu(self, mu):
        N.check(N.lib().v2h_sdp_set_mu(self._h, float(mu)))
