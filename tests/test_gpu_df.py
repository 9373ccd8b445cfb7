"""GPU parity of the density-fitted intake (v2h_batch_set_df_integrals).

The device contracts (Q|mn) into h' and (pq|rs) and packs c; the oracle
(orc_df_integrals, pinned in test_oracle_df.py) does the same with plain
strided dots.  Both sum the auxiliary index in different orders, so the
bound is 1e-13 relative to the largest |c| (and to |E_fzc|), not bit
equality.  Covered: no core, scattered core and active indices, a
symmetry-blocked instance, a 64-orbital space with 8 core orbitals and 16
active ones, two intakes into one batch, the argument checks, and one solve
from DF integrals against the oracle's solve.
"""
import numpy as np
import pytest

from hilbert_amd.batch import Batch, Problem, options
from hilbert_amd.integrals import synthetic_df
from oracle.pyoracle import OracleProblem, df_integrals

pytestmark = pytest.mark.gpu

CASES = [
    # name, nfull, core, active, na, nb, naux, active irreps (None: C1), nirrep
    ("nocore", 6, [], [5, 1, 3, 0], 2, 2, 11, None, 1),
    ("scattered", 9, [7, 2, 4], [8, 0, 5, 1, 6], 3, 2, 13, None, 1),
    ("c2v", 10, [0, 9], [1, 2, 3, 4, 5, 6], 3, 3, 20, [0, 0, 1, 1, 2, 3], 4),
    ("n16_core8", 64, list(range(8)), list(range(8, 24)), 8, 8, 160, None, 1),
]


def _full_irreps(nfull, active, irr, rng):
    """Orbital irreps over the full space that agree with the active labels."""
    full = rng.integers(0, 4, nfull)
    if irr is not None:
        full[np.asarray(active)] = irr
    return full


def _case(case, seed):
    name, nfull, core, active, na, nb, naux, irr, H = case
    rng = np.random.default_rng(seed)
    full_irr = _full_irreps(nfull, active, irr, rng) if irr is not None else None
    h, qmo, _ = synthetic_df(nfull, naux, rng, irreps=full_irr)
    return Problem(len(active), na, nb, irr, H if irr is not None else None), h, qmo


def _ref_c(case, h, qmo):
    name, nfull, core, active, na, nb, naux, irr, H = case
    ha, ga, efzc = df_integrals(h, qmo, core, active)
    P = OracleProblem(len(active), na, nb, irr if irr is not None else [0] * len(active), H)
    return P.pack_c(ha, ga), efzc


@pytest.fixture(scope="module")
def df_batch():
    made = [_case(c, 40 + k) for k, c in enumerate(CASES)]
    B = Batch([m[0] for m in made], options())
    yield B, made
    B.close()


def test_df_cost_matches_oracle(df_batch):
    B, made = df_batch
    for i, case in enumerate(CASES):
        _, h, qmo = made[i]
        efzc = B.set_df_integrals(i, h, qmo, case[2], case[3])
        ref, ref_e = _ref_c(case, h, qmo)
        got = B.c(i)
        print(case[0], "max |dc|", np.abs(got - ref).max(), "dE_fzc", efzc - ref_e, flush=True)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), case[0]
        assert abs(efzc - ref_e) <= 1e-13 * max(1.0, abs(ref_e)), case[0]
        if not case[2]:
            assert efzc == 0.0


def test_df_second_intake_replaces_the_first(df_batch):
    B, made = df_batch
    i = 1
    case = CASES[i]
    _, h, qmo = _case(case, 99)
    B.set_df_integrals(i, h, qmo, case[2], case[3])
    ref, _ = _ref_c(case, h, qmo)
    assert np.abs(B.c(i) - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_df_argument_checks(df_batch):
    B, made = df_batch
    case = CASES[1]
    _, h, qmo = made[1]
    with pytest.raises(Exception, match="twice"):
        B.set_df_integrals(1, h, qmo, [7, 2, 4], [8, 0, 5, 1, 7])   # 7 is core and active
    with pytest.raises(Exception, match="out of range"):
        B.set_df_integrals(1, h, qmo, [7, 2, 4], [8, 0, 5, 1, 9])
    with pytest.raises(ValueError):
        B.set_df_integrals(1, h, qmo[:, :-1], [7, 2, 4], case[3])
    with pytest.raises(ValueError):
        B.set_df_integrals(1, h, qmo, [7, 2, 4], case[3][:-1])


def test_df_solve_matches_oracle():
    case = ("solve", 7, [6], [0, 3, 5, 1], 2, 2, 10, None, 1)
    prob, h, qmo = _case(case, 5)
    tight = dict(e_convergence=1e-9, r_convergence=1e-8)
    B = Batch([prob], options(maxiter=50000, **tight))
    efzc = B.set_df_integrals(0, h, qmo, case[2], case[3])
    assert B.solve_to_convergence() == 1
    ha, ga, ref_e = df_integrals(h, qmo, case[2], case[3])
    P = OracleProblem(4, 2, 2)
    r = P.solve(P.pack_c(ha, ga), maxiter=50000, nthreads=4, **tight)
    assert r["converged"]
    st = B.status(0)
    assert st["converged"] and st["status_bits"] == 0
    # the bound of the other solves at these thresholds (the oracle's own
    # LAPACK / Jacobi spread is 2.2e-8 Eh)
    assert abs((st["energy_primal"] + efzc) - (r["energy_primal"] + ref_e)) < 5e-8
    B.close()
