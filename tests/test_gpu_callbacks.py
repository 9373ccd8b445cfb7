"""GPU: BPSDP with caller-supplied maps (v2h_sdp_*, the generic libsdp
contract).  With the oracle's maps as the callbacks, the solve reaches the
oracle's energy (1e-6 Eh at e_convergence 1e-7, r_convergence 1e-6) and
satisfies its own stopping criteria; y, z and mu persist across calls."""
import numpy as np
import pytest

from hilbert_amd.batch import options
from hilbert_amd.integrals import hubbard_integrals, synthetic_integrals
from hilbert_amd.sdp import BPSDPSolver
from oracle.pyoracle import OracleProblem

pytestmark = pytest.mark.gpu

E, R = 1e-7, 1e-6


def _case(name):
    if name == "hubbard6":
        h, g = hubbard_integrals(6, 1.0, 4.0)
        return OracleProblem(6, 3, 3, [0] * 6, 1), h, g
    rng = np.random.default_rng(3)
    h, g = synthetic_integrals(5, 12, rng)
    return OracleProblem(5, 3, 2, [0] * 5, 1), h, g


@pytest.mark.parametrize("name", ["hubbard6", "doublet5"])
def test_callback_solve_matches_oracle(name):
    P, h, g = _case(name)
    c = P.pack_c(h, g)
    ref = P.solve(c, e_convergence=E, r_convergence=R, maxiter=50000)
    assert ref["converged"]
    S = BPSDPSolver(P.n_primal, P.n_dual, options(e_convergence=E, r_convergence=R, maxiter=50000))
    x = P.guess()
    done = False
    calls = 0
    while not done and calls < 100:  # chunks of 500 outer iterations, state kept between calls
        done = S.solve(x, P.b(), c, P.dims, 500, P.Au, P.ATu)
        calls += 1
    st = S.status()
    assert done and st["converged"] and calls > 1
    assert abs(st["energy_primal"] - ref["energy_primal"]) < 1e-6
    assert abs(float(c @ x) - st["energy_primal"]) < 1e-10
    assert st["eps_p"] < R and st["eps_d"] < R
    assert np.abs(P.Au(x) - P.b()).max() < 1e-4
    S.close()
