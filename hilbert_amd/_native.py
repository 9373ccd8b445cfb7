"""ctypes bindings of libv2rdm_hip.so (include/v2rdm_hip.h).

The library is built in-tree by hilbert_amd.build; importing this module does
not build it.  There is no CPU fallback: a missing library is an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libv2rdm_hip.so")

V2H_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP error", 3: "out of memory", 4: "internal error"}


class V2hError(RuntimeError):
    """A non-zero status from the native library."""


class SdpOptions(C.Structure):
    _fields_ = [("sdp_objective_convergence", C.c_double), ("sdp_error_convergence", C.c_double),
                ("cg_convergence", C.c_double), ("cg_maxiter", C.c_int), ("maxiter", C.c_int),
                ("mu_update_frequency", C.c_int), ("mu0", C.c_double)]


class ProblemDesc(C.Structure):
    _fields_ = [("nmo", C.c_int), ("nalpha", C.c_int), ("nbeta", C.c_int), ("multiplicity", C.c_int),
                ("nirrep", C.c_int), ("orbital_irreps", C.POINTER(C.c_int)),
                ("constrain_q2", C.c_int), ("constrain_g2", C.c_int), ("constrain_spin", C.c_int),
                ("constrain_sz", C.c_int), ("constrain_t1", C.c_int),
                ("constrain_t2", C.c_int), ("constrain_d3", C.c_int)]


class Tuning(C.Structure):
    """v2h_tuning: launch configuration that changes no result bit."""
    _fields_ = [("class_lane", C.c_int * 7), ("fallback_lane", C.c_int * 7), ("refine_lane", C.c_int),
                ("cg_margin", C.c_int), ("xcd_remap", C.c_int), ("graphs", C.c_int), ("pipeline", C.c_int)]

    def as_dict(self):
        return {"class_lane": list(self.class_lane), "fallback_lane": list(self.fallback_lane),
                "refine_lane": self.refine_lane, "cg_margin": self.cg_margin, "xcd_remap": self.xcd_remap,
                "graphs": self.graphs, "pipeline": self.pipeline}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for k in range(7):
            t.class_lane[k] = int(d["class_lane"][k])
            t.fallback_lane[k] = int(d["fallback_lane"][k])
        t.refine_lane, t.cg_margin, t.xcd_remap = int(d["refine_lane"]), int(d["cg_margin"]), int(d["xcd_remap"])
        t.graphs = int(d.get("graphs", 0))
        t.pipeline = int(d.get("pipeline", 0))
        return t


class OrbOptOptions(C.Structure):
    _fields_ = [("g_convergence", C.c_double), ("e_convergence", C.c_double), ("maxiter", C.c_int),
                ("active_active", C.c_int)]


class OrbOptResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("nrot", C.c_int), ("evaluations", C.c_int),
                ("rotations", C.c_int), ("gnorm", C.c_double),
                ("dE", C.c_double), ("energy", C.c_double), ("efzc", C.c_double)]


class Status(C.Structure):
    _fields_ = [("mu", C.c_double), ("energy_primal", C.c_double), ("energy_dual", C.c_double),
                ("primal_error", C.c_double), ("dual_error", C.c_double),
                ("oiter_total", C.c_int64), ("iiter_total", C.c_int64), ("converged", C.c_int),
                ("active", C.c_int), ("status_bits", C.c_int), ("last_cg_iterations", C.c_int)]


# v2h_monitor_fn: (instance, oiter, iiter, E_p, E_d, mu, eps_p, eps_d, data).
MONITOR = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                      C.c_double, C.c_double, C.c_void_p)
# v2h_sdp_monitor_fn, libsdp's monitor: (print_level, oiter, iiter, E_p, E_d, mu, eps_p, eps_d, data).
SDP_MONITOR = MONITOR

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/v2rdm_hip.h
SIGNATURES = {
    "v2h_last_error": (C.c_char_p, []),
    "v2h_default_options": (C.c_int, [C.POINTER(SdpOptions)]),
    "v2h_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "v2h_calibration_copy": (C.c_int, [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "v2h_batch_create": (C.c_int, [C.POINTER(ProblemDesc), C.c_int, C.POINTER(SdpOptions), C.c_int,
                                   C.POINTER(_vp)]),
    "v2h_batch_destroy": (C.c_int, [_vp]),
    "v2h_batch_count": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "v2h_batch_set_options": (C.c_int, [_vp, C.POINTER(SdpOptions)]),
    "v2h_batch_set_cg_tolerance_factor": (C.c_int, [_vp, C.c_double]),
    "v2h_batch_sizes": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int)]),
    "v2h_batch_block_dims": (C.c_int, [_vp, C.c_int, _ip]),
    "v2h_batch_orbital_order": (C.c_int, [_vp, C.c_int, _ip]),
    "v2h_batch_set_integrals": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "v2h_batch_set_df_integrals": (C.c_int, [_vp, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _ip, _ip, _dp]),
    "v2h_batch_write_fcidump": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _dp, C.c_int, _dp, C.c_int, _ip, _ip,
                                          C.c_double, C.c_char_p, C.c_char_p, _vp, _vp]),
    "v2h_batch_optimize_orbitals": (C.c_int, [_vp, C.c_int, _vp, C.POINTER(OrbOptOptions), C.POINTER(OrbOptResult)]),
    "v2h_batch_get_df_orbitals": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "v2h_batch_orbital_gradient": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _vp, _vp, _vp, C.c_int,
                                             C.POINTER(C.c_int)]),
    "v2h_batch_set_cost": (C.c_int, [_vp, C.c_int, _dp]),
    "v2h_batch_set_guess": (C.c_int, [_vp, C.c_int, _vp]),
    "v2h_batch_set_guess_hf": (C.c_int, [_vp, C.c_int]),
    "v2h_batch_reset": (C.c_int, [_vp, C.c_int]),
    "v2h_batch_solve": (C.c_int, [_vp, C.c_int, _vp, _vp, C.POINTER(C.c_int)]),
    "v2h_batch_status": (C.c_int, [_vp, C.c_int, C.POINTER(Status)]),
    "v2h_batch_last_solve_seconds": (C.c_double, [_vp]),
    "v2h_batch_get_x": (C.c_int, [_vp, C.c_int, _dp]),
    "v2h_batch_get_y": (C.c_int, [_vp, C.c_int, _dp]),
    "v2h_batch_get_z": (C.c_int, [_vp, C.c_int, _dp]),
    "v2h_batch_set_y": (C.c_int, [_vp, C.c_int, _dp]),
    "v2h_batch_set_z": (C.c_int, [_vp, C.c_int, _dp]),
    "v2h_batch_get_mu": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double)]),
    "v2h_batch_set_mu": (C.c_int, [_vp, C.c_int, C.c_double]),
    "v2h_batch_get_c": (C.c_int, [_vp, C.c_int, _dp]),
    "v2h_batch_get_b": (C.c_int, [_vp, C.c_int, _dp]),
    "v2h_batch_get_opdm": (C.c_int, [_vp, C.c_int, C.c_char_p, _dp]),
    "v2h_batch_get_tpdm": (C.c_int, [_vp, C.c_int, _dp]),
    "v2h_batch_get_opdm_sparse": (C.c_int, [_vp, C.c_int, C.c_char_p, _vp, _vp, C.c_int64,
                                            C.POINTER(C.c_int64)]),
    "v2h_batch_get_tpdm_sparse": (C.c_int, [_vp, C.c_int, C.c_char_p, _vp, _vp, C.c_int64,
                                            C.POINTER(C.c_int64)]),
    "v2h_batch_Au": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "v2h_batch_ATu": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "v2h_batch_cost_adjoint": (C.c_int, [_vp, C.c_int, _vp, _dp, _dp]),
    "v2h_batch_set_profiling": (C.c_int, [_vp, C.c_int]),
    "v2h_batch_set_graphs": (C.c_int, [_vp, C.c_int]),
    "v2h_batch_get_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "v2h_batch_set_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "v2h_batch_kernel_times": (C.c_int, [_vp, _dp, _lp, _dp, _lp]),
    "v2h_batch_reset_kernel_times": (C.c_int, [_vp]),
    "v2h_batch_launch_stats": (C.c_int, [_vp, _lp]),
    "v2h_batch_set_projection": (C.c_int, [_vp, C.c_int]),
    "v2h_batch_Au_noalloc": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "v2h_batch_projection_stats": (C.c_int, [_vp, _lp]),
    "v2h_batch_set_refine_params": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_int]),
    "v2h_sdp_create": (C.c_int, [C.c_int64, C.c_int64, C.POINTER(SdpOptions), C.c_int, C.POINTER(_vp)]),
    "v2h_sdp_destroy": (C.c_int, [_vp]),
    "v2h_sdp_set_options": (C.c_int, [_vp, C.POINTER(SdpOptions)]),
    "v2h_sdp_solve": (C.c_int, [_vp, _dp, _dp, _dp, _ip, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp,
                                C.POINTER(C.c_int)]),
    "v2h_sdp_status": (C.c_int, [_vp, C.POINTER(Status)]),
    "v2h_sdp_get_y": (C.c_int, [_vp, _dp]),
    "v2h_sdp_get_z": (C.c_int, [_vp, _dp]),
    "v2h_sdp_set_y": (C.c_int, [_vp, _dp]),
    "v2h_sdp_set_z": (C.c_int, [_vp, _dp]),
    "v2h_sdp_get_mu": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "v2h_sdp_set_mu": (C.c_int, [_vp, C.c_double]),
    "v2h_structure_create": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(_vp)]),
    "v2h_structure_destroy": (C.c_int, [_vp]),
    "v2h_structure_sizes": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "v2h_structure_block_dims": (C.c_int, [_vp, _ip]),
    "v2h_structure_tables": (C.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "v2h_structure_constraints": (C.c_int, [_vp, _lp, _ip, _dp, _dp]),
    "v2h_structure_cost_map": (C.c_int, [_vp, _ip, _ip]),
    "v2h_structure_gather_stats": (C.c_int, [_vp, _lp]),
    "v2h_structure_opdm": (C.c_int, [_vp, _dp, C.c_char_p, _dp]),
    "v2h_structure_tpdm": (C.c_int, [_vp, _dp, _dp]),
    "v2h_structure_hf_guess": (C.c_int, [_vp, _dp]),
    "v2h_structure_write_rdm_files": (C.c_int, [_vp, _dp, C.c_int, _ip, _ip, C.c_char_p, C.c_int, _lp]),
    "v2h_batch_write_rdm_files": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, C.c_char_p, C.c_int, _lp]),
}

_lib = None


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing: run `python -m hilbert_amd.build` "
                              "(the solver has no CPU fallback)")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != V2H_OK:
        msg = lib().v2h_last_error().decode(errors="replace")
        raise V2hError(f"{ERRORS.get(rc, rc)}: {msg}")


RDM_FILE_COUNTS = ("d1a", "d1b", "d2aa", "d2bb", "d2ab", "d2", "aa.iwl", "bb.iwl", "ab.iwl")
# which files: hilbert's options that call each writer (v2rdm_solver.cc:1474-1509)
RDM_TPDM_ACTIVE, RDM_TPDM_FULL, RDM_TPDM_SPIN_FREE, RDM_OPDM_FULL, RDM_TPDM_IWL = 1, 2, 4, 8, 16
RDM_ALL_FULL = RDM_TPDM_FULL | RDM_TPDM_SPIN_FREE | RDM_OPDM_FULL | RDM_TPDM_IWL


def write_rdm_files(fn, head, prefix, active, core, which):
    """Call a *_write_rdm_files entry point; returns {file suffix: records}
    for the files written."""
    core = np.ascontiguousarray(core if core is not None else [], dtype=np.int32)
    active = np.ascontiguousarray(active, dtype=np.int32)
    counts = np.zeros(len(RDM_FILE_COUNTS), dtype=np.int64)
    core_arg = core if len(core) else np.zeros(1, dtype=np.int32)
    check(fn(*head, len(core), core_arg, active, os.fsencode(prefix), int(which), counts))
    return {k: int(v) for k, v in zip(RDM_FILE_COUNTS, counts) if v >= 0}


def make_desc(nmo, nalpha, nbeta, irreps=None, nirrep=None, multiplicity=0, q2=True, g2=True, spin=True,
              sz=True, t1=False, t2=False, d3=False):
    """ProblemDesc plus the array that keeps its irrep pointer alive."""
    arr = None
    ptr = None
    if irreps is not None:
        arr = (C.c_int * nmo)(*[int(v) for v in irreps])
        ptr = C.cast(arr, C.POINTER(C.c_int))
        if nirrep is None:
            nirrep = 1
            while nirrep <= max(int(v) for v in irreps):
                nirrep *= 2
    d = ProblemDesc(int(nmo), int(nalpha), int(nbeta), int(multiplicity), int(nirrep or 1), ptr,
                    int(q2), int(g2), int(spin), int(sz), int(t1), int(t2), int(d3))
    return d, arr


def default_options() -> SdpOptions:
    o = SdpOptions()
    check(lib().v2h_default_options(C.byref(o)))
    return o


def calibration_copy(n: int, repeats: int = 3, device: int = 0) -> float:
    """Mean milliseconds of one plain device copy of n doubles (k_copy)."""
    ms = C.c_double(0.0)
    check(lib().v2h_calibration_copy(device, n, repeats, C.byref(ms)))
    return ms.value


def device_count() -> int:
    c = C.c_int(0)
    rc = lib().v2h_device_count(C.byref(c))
    return c.value if rc == V2H_OK else 0
