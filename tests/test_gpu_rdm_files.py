"""GPU: hilbert's RDM files (write_opdm.cc:40, write_tpdm.cc:40-595,
write_tpdm_iwl.cc:45) of solved instances, through v2h_batch_write_rdm_files.

* the files of a solved instance equal the oracle's records of the same
  primal vector, bit for bit (oracle/writers.py, pinned on the CPU by the
  brute-force energy of embedded states);
* with no core and the identity labels, the PSIO records and the IWL values
  reproduce the solver's energy c.x from the user's (h, g), also for orbitals
  whose irreps are not sorted (the labels go through the Pitzer-to-user map).
"""
import os

import numpy as np
import pytest

from hilbert_amd import configs
from hilbert_amd.batch import Batch, Problem, options
from hilbert_amd.integrals import synthetic_integrals
from oracle import writers as W
from oracle.pyoracle import OracleProblem

pytestmark = pytest.mark.gpu

OPTS = dict(e_convergence=1e-7, r_convergence=1e-6, maxiter=40000)


def _energy_from_files(prefix, h, g):
    e = 0.0
    for ext, key in (("d1a", "D1a"), ("d1b", "D1b")):
        r = np.frombuffer(W.read_keyed(prefix + "." + ext)[key], dtype=W.OPDM)
        e += float(np.sum(h[r["i"], r["j"]] * r["value"]))
    e_iwl = e
    for ext, key, w in (("d2aa", "D2aa", 0.5), ("d2bb", "D2bb", 0.5), ("d2ab", "D2ab", 1.0)):
        t = np.frombuffer(W.read_keyed(prefix + "." + ext)[key], dtype=W.TPDM)
        e += w * float(np.sum(g[t["i"], t["k"], t["j"], t["l"]] * t["value"]))
    for ext, w in (("aa.iwl", 2.0), ("bb.iwl", 2.0), ("ab.iwl", 1.0)):
        q, _ = W.read_iwl(prefix + "." + ext)
        e_iwl += w * float(np.sum(g[q["i"], q["j"], q["k"], q["l"]] * q["value"]))
    return e, e_iwl


def test_files_of_a_solved_instance_equal_the_oracle_records(tmp_path):
    irr = [0, 0, 1, 1, 2, 3]
    prob = Problem(6, 3, 3, irr, 4)
    h, g = synthetic_integrals(6, 18, np.random.default_rng(611), irreps=np.array(irr))
    B = Batch([prob], options(**OPTS))
    B.set_integrals(0, h, g)
    B.solve_to_convergence()
    assert B.status(0)["converged"]
    x = B.x(0)
    core, active = [0, 8], [6, 5, 4, 3, 2, 1]
    prefix = os.path.join(str(tmp_path), "rdm")
    counts = B.write_rdm_files(0, prefix, active, core)
    B.close()
    P = OracleProblem(6, 3, 3, irr, 4)
    for beta, ext, key in ((False, "d1a", "D1a"), (True, "d1b", "D1b")):
        ref = W.opdm_records(P, x, active, core, beta)
        assert W.read_keyed(prefix + "." + ext)[key] == ref.tobytes()
        assert counts[ext] == len(ref)
    for part, ext, key in ((0, "d2aa", "D2aa"), (1, "d2bb", "D2bb"), (2, "d2ab", "D2ab")):
        ref = W.tpdm_records(P, x, active, core, part)
        assert W.read_keyed(prefix + "." + ext)[key] == ref.tobytes()
        assert counts[ext] == len(ref)
    ref = W.spin_free_records(P, x, active, core)
    assert W.read_keyed(prefix + ".d2")["D2"] == ref.tobytes()
    for part, ext in ((0, "aa.iwl"), (1, "bb.iwl"), (2, "ab.iwl")):
        ref = W.iwl_records(P, x, active, core, part)
        got, _ = W.read_iwl(prefix + "." + ext)
        assert got.tobytes() == ref[np.abs(ref["value"]) > W.IWL_CUTOFF].tobytes()


def _unsorted():
    irr = [1, 0, 3, 0, 2, 1, 3]
    h, g = synthetic_integrals(7, 21, np.random.default_rng(712), irreps=np.array(irr))
    return Problem(7, 4, 3, irr, 4), h, g


@pytest.mark.parametrize("which", ["c1_10", "c2v_7_unsorted"])
def test_files_reproduce_the_solver_energy(which, tmp_path):
    prob, h, g = configs.c1_10()[0] if which == "c1_10" else _unsorted()
    B = Batch([prob], options(**OPTS))
    B.set_integrals(0, h, g)
    B.solve_to_convergence()
    st = B.status(0)
    assert st["converged"]
    prefix = os.path.join(str(tmp_path), "rdm")
    B.write_rdm_files(0, prefix, list(range(prob.nmo)))
    B.close()
    e_psio, e_iwl = _energy_from_files(prefix, h, g)
    e = st["energy_primal"]
    assert abs(e_psio - e) < 1e-10 * max(1.0, abs(e)), (e_psio, e)
    assert abs(e_iwl - e) < 1e-10 * max(1.0, abs(e)), (e_iwl, e)
