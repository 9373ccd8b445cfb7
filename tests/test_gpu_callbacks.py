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


def test_monitor_gets_libsdp_arguments():
    """hilbert's bpsdp_monitor prints only for print_level > 0
    (v2rdm_solver.cc:115-124) and receives the print_level given to solve
    (v2rdm_solver.cc:1340): a monitor gated the same way logs one line per
    outer iteration at print_level 1 and none at print_level 0."""
    P, h, g = _case("hubbard6")
    c = P.pack_c(h, g)
    S = BPSDPSolver(P.n_primal, P.n_dual, options(maxiter=50000))
    lines, seen = [], []

    def bpsdp_monitor(print_level, oiter, iiter, ep, ed, mu, eps_p, eps_d, data):
        seen.append((print_level, data))
        if print_level > 0:
            lines.append((oiter, iiter, ep, ed, mu, eps_p, eps_d))

    x = P.guess()
    S.solve(x, P.b(), c, P.dims, 7, P.Au, P.ATu, monitor=bpsdp_monitor, print_level=1, data="solver")
    assert [l[0] for l in lines] == list(range(1, 8))
    assert all(s == (1, "solver") for s in seen)
    st = S.status()
    oi, ii, ep, ed, mu, eps_p, eps_d = lines[-1]
    assert (ep, ed, mu, eps_p, eps_d) == (st["energy_primal"], st["energy_dual"], st["mu"], st["eps_p"], st["eps_d"])
    assert sum(l[1] for l in lines) == st["iiter"] and ii >= 0
    n = len(lines)
    S.solve(x, P.b(), c, P.dims, 5, P.Au, P.ATu, monitor=bpsdp_monitor, print_level=0)
    assert len(lines) == n and len(seen) == 12 and seen[-1][0] == 0
    S.close()
