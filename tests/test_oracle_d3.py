"""Pin the oracle's D3 condition (CONSTRAIN_D3; d3.cc:45-406, blocks
v2rdm_solver.cc:4000-4026, rows :3624-3675).

* exact-state feasibility: for brute-force spin-adapted states of definite
  symmetry, the DQG primal plus the exact 3-RDM blocks
  D3(ijk,lmn) = <a+_i a+_j a+_k a_n a_m a_l> (spins aaa, bbb, aab, bba)
  satisfy every row (contractions with and without CONSTRAIN_SZ, and the
  singlet spin rows) to 1e-12; the other operator order and a foreign D3
  violate them;
* exact D3 blocks are positive semidefinite;
* sizes and adjointness.  The bound E_DQG <= E_DQG+D3 <= E_FCI is checked
  on the GPU (tests/test_gpu_t1t2.py), where the solves are fast: with
  fewer than three electrons of a spin the D3aaa / D3bbb blocks are forced
  to zero (a face with no interior point), on which the boundary-point
  iteration converges slowly.
"""
import numpy as np
import pytest

from fci import Fock, RDMs, pack_d3, pack_primal, spin_adapted_random_state
from oracle.pyoracle import OracleProblem

CASES = [
    (4, 2, 2, [0, 0, 0, 0], 1),
    (4, 2, 1, [0, 0, 0, 0], 1),
    (5, 3, 3, [0, 0, 1, 1, 1], 2),
    (5, 3, 2, [0, 1, 1, 2, 3], 4),
]
_FOCK = {}


def _fock(n):
    if n not in _FOCK:
        _FOCK[n] = Fock(n)
    return _FOCK[n]


def _exact_x(P, case, sz, seed, order="nml"):
    n, na, nb, irr, H = case
    F = _fock(n)
    psi = spin_adapted_random_state(F, na, nb, np.random.default_rng(seed), irreps=irr if H > 1 else None, gamma=0)
    P0 = OracleProblem(n, na, nb, irr, H, sz=sz)
    x = np.zeros(P.n_primal)
    x[:P0.n_primal] = pack_primal(P0, RDMs(F, psi))
    return pack_d3(P, F, psi, x, order), P0


@pytest.mark.parametrize("sz", [True, False])
@pytest.mark.parametrize("case", CASES)
def test_exact_state_satisfies_d3_rows(case, sz):
    n, na, nb, irr, H = case
    P = OracleProblem(n, na, nb, irr, H, sz=sz, d3=True)
    x, P0 = _exact_x(P, case, sz, 41)
    res = P.Au(x) - P.b()
    assert np.abs(res).max() < 1e-12, (np.abs(res[:P0.n_dual]).max(), np.abs(res[P0.n_dual:]).max())
    off = P.t1_offsets()
    taab, taaa = P.trip_counts()
    for name, t in (("d3aaa", taaa), ("d3bbb", taaa), ("d3aab", taab), ("d3bba", taab)):
        for h in range(H):
            d = int(t[h])
            if d == 0:
                continue
            blk = x[off[name][h]:off[name][h] + d * d].reshape(d, d)
            assert np.abs(blk - blk.T).max() < 1e-12
            assert np.linalg.eigvalsh(blk).min() > -1e-12
    # the other operator order, <a+_i a+_j a+_k a_l a_m a_n>, breaks the rows
    y, _ = _exact_x(P, case, sz, 41, order="lmn")
    assert np.abs(P.Au(y)[P0.n_dual:]).max() > 1e-2


def test_foreign_d3_violates_rows():
    case = CASES[0]
    n, na, nb, irr, H = case
    P = OracleProblem(n, na, nb, irr, H, d3=True)
    x, P0 = _exact_x(P, case, True, 41)
    y, _ = _exact_x(P, case, True, 42)
    x[P0.n_primal:] = y[P0.n_primal:]
    res = P.Au(x) - P.b()
    assert np.abs(res[:P0.n_dual]).max() < 1e-12 and np.abs(res[P0.n_dual:]).max() > 1e-3


@pytest.mark.parametrize("case", CASES)
def test_d3_sizes(case):
    n, na, nb, irr, H = case
    for sz in (True, False):
        P = OracleProblem(n, na, nb, irr, H, sz=sz, d3=True)
        P0 = OracleProblem(n, na, nb, irr, H, sz=sz)
        taab, taaa = P.trip_counts()
        T = P.tables()
        gab, gaa = T["gems_ab"].astype(np.int64), T["gems_aa"].astype(np.int64)
        rows = (4 * (gaa ** 2).sum() + 2 * (gab ** 2).sum()) if sz else (2 * (gaa ** 2).sum() + (gab ** 2).sum())
        if na == nb:
            rows += (taab.astype(np.int64) ** 2).sum() + 2 * (taaa.astype(np.int64) ** 2).sum()
        assert P.n_dual - P0.n_dual == rows
        assert P.n_primal - P0.n_primal == 2 * int((taaa.astype(np.int64) ** 2).sum() + (taab.astype(np.int64) ** 2).sum())


@pytest.mark.parametrize("case", CASES[:3])
def test_d3_adjointness(case):
    n, na, nb, irr, H = case
    P = OracleProblem(n, na, nb, irr, H, t1=True, t2=True, d3=True)
    rng = np.random.default_rng(7)
    u = rng.standard_normal(P.n_primal)
    y = rng.standard_normal(P.n_dual)
    lhs, rhs = y @ P.Au(u), P.ATu(y) @ u
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))

