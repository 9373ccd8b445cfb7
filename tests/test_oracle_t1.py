"""Pin the oracle's T1 condition (POSITIVITY DQGT1; t1.cc, triplets basis.cc:642-744).

* exact-state feasibility: for brute-force spin-adapted states of definite
  symmetry, the DQG primal plus the exact T1 blocks
  T1(ijk,lmn) = <a+_i a+_j a+_k a_n a_m a_l> + <a_i a_j a_k a+_n a+_m a+_l>
  satisfy every row, T1 rows included (b = 0 there), to 1e-12; a sign or
  index error in a T1 term leaves a residual of order 0.1;
* those exact T1 blocks are positive semidefinite;
* sizes: trip_aab = #{(i<j, k)}, trip_aaa = #{i<j<k} per irrep, T1 adds
  2 (taaa^2 + taab^2) primal elements and as many rows;
* adjointness of Au and A^T u with T1.
"""
import numpy as np
import pytest

from fci import Fock, RDMs, pack_primal, pack_t1, spin_adapted_random_state
from oracle.pyoracle import OracleProblem

CASES = [
    (4, 2, 1, [0, 0, 0, 0], 1),
    (4, 2, 2, [0, 0, 0, 0], 1),
    (5, 2, 2, [0, 0, 1, 1, 1], 2),
    (4, 2, 2, [0, 1, 2, 3], 4),
    (5, 3, 2, [0, 1, 1, 2, 3], 4),
]
_FOCK = {}


def _fock(n):
    if n not in _FOCK:
        _FOCK[n] = Fock(n)
    return _FOCK[n]


def _exact_x(P, case, seed):
    n, na, nb, irr, H = case
    F = _fock(n)
    psi = spin_adapted_random_state(F, na, nb, np.random.default_rng(seed), irreps=irr if H > 1 else None, gamma=0)
    P0 = OracleProblem(n, na, nb, irr, H)
    x = np.zeros(P.n_primal)
    x[:P0.n_primal] = pack_primal(P0, RDMs(F, psi))
    return pack_t1(P, F, psi, x), P0


@pytest.mark.parametrize("case", CASES)
def test_exact_state_satisfies_t1_rows(case):
    n, na, nb, irr, H = case
    P = OracleProblem(n, na, nb, irr, H, t1=True)
    x, P0 = _exact_x(P, case, 21)
    res = P.Au(x) - P.b()
    assert P.n_dual > P0.n_dual
    assert np.abs(res).max() < 1e-12, (np.abs(res[:P0.n_dual]).max(), np.abs(res[P0.n_dual:]).max())
    # the T1 blocks of an exact state are PSD
    off = P.t1_offsets()
    taab, taaa = P.trip_counts()
    for name, t in (("t1aaa", taaa), ("t1bbb", taaa), ("t1aab", taab), ("t1bba", taab)):
        for h in range(H):
            d = int(t[h])
            if d == 0:
                continue
            blk = x[off[name][h]:off[name][h] + d * d].reshape(d, d)
            assert np.abs(blk - blk.T).max() < 1e-12
            assert np.linalg.eigvalsh(blk).min() > -1e-12


def test_wrong_state_violates_t1_rows():
    """A T1 block of another state breaks the rows that tie it to D2, Q2, G2."""
    case = CASES[1]
    n, na, nb, irr, H = case
    P = OracleProblem(n, na, nb, irr, H, t1=True)
    x, P0 = _exact_x(P, case, 21)
    y, _ = _exact_x(P, case, 22)
    x[P0.n_primal:] = y[P0.n_primal:]
    res = P.Au(x) - P.b()
    assert np.abs(res[:P0.n_dual]).max() < 1e-12
    assert np.abs(res[P0.n_dual:]).max() > 1e-3


@pytest.mark.parametrize("case", CASES)
def test_t1_sizes(case):
    n, na, nb, irr, H = case
    P = OracleProblem(n, na, nb, irr, H, t1=True)
    P0 = OracleProblem(n, na, nb, irr, H)
    taab, taaa = P.trip_counts()
    irr = np.asarray(irr)
    for h in range(H):
        trips = [(i, j, k) for i in range(n) for j in range(n) for k in range(n) if irr[i] ^ irr[j] ^ irr[k] == h]
        assert taab[h] == sum(1 for t in trips if t[0] < t[1])
        assert taaa[h] == sum(1 for t in trips if t[0] < t[1] < t[2])
    extra = 2 * int((taaa.astype(np.int64) ** 2).sum() + (taab.astype(np.int64) ** 2).sum())
    assert P.n_primal - P0.n_primal == extra and P.n_dual - P0.n_dual == extra
    assert P.dims.tolist() == P0.dims.tolist() + list(taaa) + list(taaa) + list(taab) + list(taab)


@pytest.mark.parametrize("case", CASES[:3])
def test_t1_adjointness(case):
    n, na, nb, irr, H = case
    P = OracleProblem(n, na, nb, irr, H, t1=True)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(P.n_primal)
    y = rng.standard_normal(P.n_dual)
    lhs, rhs = y @ P.Au(u), P.ATu(y) @ u
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_t1_bound_lies_between_dqg_and_fci():
    """Adding T1 tightens the lower bound: E_DQG <= E_DQGT1 <= E_FCI, and by
    more than the stopping band on this instance."""
    from fci import ground_state
    from hilbert_amd.integrals import synthetic_integrals
    n, na, nb = 4, 2, 2
    h, g = synthetic_integrals(n, 8, 5)
    F = _fock(n)
    e_fci, _ = ground_state(F, F.hamiltonian(h, g), na, nb)
    kw = dict(maxiter=50000, nthreads=2, e_convergence=1e-6, r_convergence=1e-5)
    P0 = OracleProblem(n, na, nb)
    r0 = P0.solve(P0.pack_c(h, g), **kw)
    P1 = OracleProblem(n, na, nb, t1=True)
    r1 = P1.solve(P1.pack_c(h, g), **kw)
    assert r0["converged"] and r1["converged"]
    e0, e1 = r0["energy_primal"], r1["energy_primal"]
    # bounds within the stopping band (5e-5 Eh); here T1 closes 90 % of the
    # DQG gap (-7.56472 / -7.56102 / -7.56067)
    assert e0 <= e1 + 5e-5 and e1 <= e_fci + 5e-5, (e0, e1, e_fci)
    assert e1 - e0 > 1e-3
