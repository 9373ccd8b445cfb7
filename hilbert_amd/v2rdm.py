"""hilbert's fixed-orbital v2RDM entry point on the MI355X solver.

``v2RDMHelper(nalpha, nbeta, nmo, h, g, options)`` mirrors the pybind class
``hilbert.v2RDMHelper`` for externally defined Hamiltonians
(python_api/python_helpers.h:160-184, python_helpers.cc:104-113, 356-399):

    helper = v2RDMHelper(nalpha, nbeta, nmo, h, g, {"E_CONVERGENCE": 1e-6})
    energy = helper.compute_energy()
    D1 = helper.get_opdm()            # spin-summed, n x n
    D2 = helper.get_tpdm()            # spin-free, (n^2, n^2): [i n + k, j n + l]
    lst = helper.get_tpdm_sparse("AB")

``h`` is n x n and ``g`` is (pq|rs) in chemists' order, flat or shaped: the
layout of the reference's external constructor (v2rdm_solver.cc:147-195).
``v2RDMHelper.hubbard(options)`` builds the Hubbard ring from the HUBBARD_*
options (v2rdm_solver.cc:839-962).

Option names and defaults are the plugin's (plugin.cc:94-100, 205-309).  Two
options are additions: IRREPS (one irrep label per orbital, caller order; the
reference's external path is C1 only, v2rdm_solver.cc:988) and NIRREP.
Settings this solver does not implement raise ``NotImplementedError`` naming
the option instead of being ignored.
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass

import numpy as np

from .batch import Batch, Problem, options as sdp_options
from .integrals import hubbard_integrals

# name -> default (plugin.cc)
DEFAULTS = {
    "E_CONVERGENCE": 1e-6,          # plugin.cc:94
    "R_CONVERGENCE": 1e-5,          # plugin.cc:97
    "MAXITER": 50,                  # plugin.cc:100
    "CG_CONVERGENCE": 1e-9,         # plugin.cc:238
    "CG_MAXITER": 10000,            # plugin.cc:241
    "MU_UPDATE_FREQUENCY": 1000,    # plugin.cc:217
    "POSITIVITY": "DQG",            # plugin.cc:220
    "CONSTRAIN_SPIN": True,         # plugin.cc:306
    "CONSTRAIN_SZ": True,           # plugin.cc:309
    "TPDM_GUESS": "RANDOM",         # plugin.cc:205
    "SDP_SOLVER": "BPSDP",          # plugin.cc:246
    "TAU_PARAMETER": 1.0,           # plugin.cc:300
    "GPC_CONSTRAINTS": "NONE",      # plugin.cc:223
    "CONSTRAIN_D3": False,          # plugin.cc:226
    "CONSTRAIN_E3": False,
    "CONSTRAIN_F3": False,
    "CONSTRAIN_Q3": False,
    "CONSTRAIN_D4": False,          # plugin.cc:303
    "HUBBARD_HAMILTONIAN": False,   # plugin.cc:249
    "HUBBARD_T": 1.0,               # plugin.cc:252
    "HUBBARD_U": 1.0,               # plugin.cc:255
    "N_HUBBARD_SITES": 4,           # plugin.cc:258
    "N_HUBBARD_SPINS": 4,           # plugin.cc:261
    "HUBBARD_MULTIPLICITY": 1,      # plugin.cc:264
    "IRREPS": None,                 # addition: per-orbital irrep labels
    "NIRREP": None,                 # addition: 1, 2, 4 or 8 (default: from IRREPS)
}

# POSITIVITY -> (q2, g2, t1, t2) (v2rdm_solver.cc:4081-4112)
_POSITIVITY = {"D": (False, False, False, False), "DQ": (True, False, False, False),
               "DG": (False, True, False, False), "DQG": (True, True, False, False),
               "DQGT1": (True, True, True, False), "DQGT2": (True, True, False, True),
               "DQGT1T2": (True, True, True, True)}


def resolve_options(options=None) -> dict:
    """Upper-cased options merged over the defaults, validated."""
    opts = dict(DEFAULTS)
    for k, v in (options or {}).items():
        key = k.upper()
        if key not in DEFAULTS:
            raise KeyError(f"unknown option {k!r}")
        opts[key] = v.upper() if isinstance(v, str) else v
    if opts["POSITIVITY"] not in _POSITIVITY:
        raise NotImplementedError(f"POSITIVITY {opts['POSITIVITY']}: only D, DQ, DG, DQG, DQGT1, DQGT2 and "
                                  "DQGT1T2 are implemented")
    if opts["SDP_SOLVER"] != "BPSDP":
        raise NotImplementedError(f"SDP_SOLVER {opts['SDP_SOLVER']}: only BPSDP is implemented")
    if opts["TPDM_GUESS"] not in ("RANDOM", "HF"):
        raise ValueError(f"TPDM_GUESS {opts['TPDM_GUESS']}: RANDOM or HF")
    if opts["GPC_CONSTRAINTS"] != "NONE":
        raise NotImplementedError("GPC_CONSTRAINTS are not implemented")
    for k in ("CONSTRAIN_E3", "CONSTRAIN_F3", "CONSTRAIN_Q3", "CONSTRAIN_D4"):
        if opts[k]:
            raise NotImplementedError(f"{k} is not implemented")
    if float(opts["TAU_PARAMETER"]) != 1.0:
        raise NotImplementedError("TAU_PARAMETER other than 1.0 is not implemented")
    return opts


@dataclass
class opdm:
    """One 1-RDM element (v2rdm_solver.h:61-65)."""
    i: int
    j: int
    value: float


@dataclass
class tpdm:
    """One 2-RDM element (v2rdm_solver.h:53-59)."""
    i: int
    j: int
    k: int
    l: int
    value: float


class v2RDMHelper:
    def __init__(self, nalpha, nbeta, nmo, h, g, options=None):
        self.options = resolve_options(options)
        o = self.options
        n = int(nmo)
        self.h = np.asarray(h, dtype=np.float64).reshape(n, n)
        self.g = np.asarray(g, dtype=np.float64).reshape(n, n, n, n)
        q2, g2, t1, t2 = _POSITIVITY[o["POSITIVITY"]]
        irreps = o["IRREPS"]
        nirrep = o["NIRREP"]
        if irreps is not None and nirrep is None:
            nirrep = 1 << int(np.ceil(np.log2(max(int(np.max(irreps)) + 1, 1))))
        self.problem = Problem(n, int(nalpha), int(nbeta), None if irreps is None else list(irreps), nirrep,
                               0, q2, g2, bool(o["CONSTRAIN_SPIN"]), bool(o["CONSTRAIN_SZ"]), t1, t2,
                               bool(o["CONSTRAIN_D3"]))
        self._batch = None
        self.energy = None

    @classmethod
    def hubbard(cls, options=None):
        """The periodic Hubbard ring of the HUBBARD_* options
        (initialize_with_hubbard_hamiltonian, v2rdm_solver.cc:839-962)."""
        o = resolve_options(options)
        nspin = int(o["N_HUBBARD_SPINS"])
        if nspin % 2 != 0:
            raise ValueError("Hubbard model only works for even numbers of electrons currently.")
        n = int(o["N_HUBBARD_SITES"])
        h, g = hubbard_integrals(n, float(o["HUBBARD_T"]), float(o["HUBBARD_U"]))
        return cls(nspin // 2, nspin // 2, n, h, g, options)

    def _sdp_options(self):
        o = self.options
        return sdp_options(e_convergence=float(o["E_CONVERGENCE"]), r_convergence=float(o["R_CONVERGENCE"]),
                           cg_convergence=float(o["CG_CONVERGENCE"]), cg_maxiter=int(o["CG_MAXITER"]),
                           maxiter=int(o["MAXITER"]), mu_update_frequency=int(o["MU_UPDATE_FREQUENCY"]))

    def compute_energy(self) -> float:
        """Solve to convergence (or MAXITER outer iterations, in chunks of
        MU_UPDATE_FREQUENCY as v2rdm_solver.cc:1298-1361) and return c.x
        (no nuclear or frozen-core term for external Hamiltonians)."""
        if self._batch is None:
            self._batch = Batch([self.problem], self._sdp_options())
            self._batch.set_integrals(0, self.h, self.g)
            if self.options["TPDM_GUESS"] == "HF":
                self._batch.set_guess_hf(0)
        self._batch.solve_to_convergence()
        st = self._batch.status(0)
        self.status = st
        self.energy = st["energy_primal"]
        if not st["converged"] or st["status_bits"]:
            warnings.warn("v2RDM SDP did not converge (%d outer iterations, MAXITER %d, status bits %d): "
                          "the energy is that of the last iterate" % (st["oiter"], int(self.options["MAXITER"]),
                                                                      st["status_bits"]),
                          RuntimeWarning, stacklevel=2)
        return self.energy

    def _need(self):
        if self._batch is None:
            raise RuntimeError("compute_energy() has not been called")
        return self._batch

    def get_opdm(self):
        """Spin-summed 1-RDM, n x n in the caller's orbital order (get_opdm,
        python_api.cc:374-394)."""
        return self._need().opdm(0, "SUM")

    def get_tpdm(self):
        """Spin-free 2-RDM as the (n^2, n^2) matrix [i n + k, j n + l]
        (get_tpdm, python_api.cc:396-456)."""
        n = self.problem.nmo
        return self._need().tpdm(0).reshape(n * n, n * n)

    def get_opdm_sparse(self, type):
        """List of opdm(i, j, value), type "A", "B" or "SUM" (python_api.cc:39-113)."""
        idx, val = self._need().opdm_sparse(0, type.upper())
        return [opdm(int(a), int(b), float(v)) for (a, b), v in zip(idx, val)]

    def get_tpdm_sparse(self, type):
        """List of tpdm(i, j, k, l, value), type "AA", "BB", "AB", "BA" or "SUM"
        (python_api.cc:115-371)."""
        idx, val = self._need().tpdm_sparse(0, type.upper())
        return [tpdm(int(a), int(b), int(c), int(d), float(v)) for (a, b, c, d), v in zip(idx, val)]

    def x(self):
        """The converged primal vector (layout of v2rdm_solver.cc:3729-3846)."""
        return self._need().x(0)

    def close(self):
        if self._batch is not None:
            self._batch.close()
            self._batch = None
