"""RDM export (v2RDMSolver::get_opdm / get_tpdm, v2rdm_casscf/python_api.cc:374-456)
from a primal vector, host side (no GPU).

* exact RDMs of brute-force wavefunctions packed into the primal layout come
  back as the spin-summed 1-RDM and the spin-free 2-RDM
  Gamma[i, k, j, l] = sum_{s,t} <a+_is a+_jt a_lt a_ks>, with and without
  point-group symmetry (the reference exports the dense 2-RDM in C1 only);
* in C1 the dense 2-RDM equals the oracle's restatement of the same loop on
  arbitrary (non-physical) vectors, element by element.
"""
import numpy as np
import pytest

from fci import Fock, RDMs, pack_primal, spin_adapted_random_state
from hilbert_amd.structure import Structure
from oracle.pyoracle import OracleProblem

CASES = [
    (4, 2, 2, None, 1),
    (5, 3, 2, None, 1),
    (5, 2, 2, [0, 0, 1, 2, 3], 4),
    (6, 3, 3, [0, 0, 1, 1, 2, 3], 4),
    (6, 3, 2, [0, 1, 2, 3, 5, 6], 8),
]


def _spin_free(R, n):
    m = R.m
    G = np.zeros((n, n, n, n))
    for s in range(2):
        for t in range(2):
            for i in range(n):
                for j in range(n):
                    for k in range(n):
                        for l in range(n):
                            G[i, k, j, l] += R.D2[m(i, s), m(j, t), m(l, t), m(k, s)]
    return G


@pytest.mark.parametrize("n,na,nb,irr,H", CASES)
def test_exact_rdms_round_trip(n, na, nb, irr, H):
    rng = np.random.default_rng(n * 7 + na + 11 * H)
    irreps = irr if irr is not None else [0] * n
    P = OracleProblem(n, na, nb, irreps, H)
    S = Structure(n, na, nb, irr, H)
    F = Fock(n)
    psi = spin_adapted_random_state(F, na, nb, rng, irreps=irreps)
    R = RDMs(F, psi)
    x = pack_primal(P, R)
    m = R.m
    d1a = np.array([[R.D1[m(p, 0), m(q, 0)] for q in range(n)] for p in range(n)])
    d1b = np.array([[R.D1[m(p, 1), m(q, 1)] for q in range(n)] for p in range(n)])
    # symmetry-forbidden elements are not stored
    same = np.equal.outer(np.asarray(irreps), np.asarray(irreps))
    assert np.abs(S.opdm(x, "A") - d1a * same).max() < 1e-12
    assert np.abs(S.opdm(x, "B") - d1b * same).max() < 1e-12
    assert np.abs(S.opdm(x, "SUM") - (d1a + d1b) * same).max() < 1e-12
    assert np.abs(S.tpdm(x) - _spin_free(R, n)).max() < 1e-12
    # trace: sum_ij Gamma[i, i, j, j] = N (N - 1)
    N = na + nb
    assert abs(np.einsum("iijj->", S.tpdm(x)) - N * (N - 1)) < 1e-10


@pytest.mark.parametrize("n", [3, 4, 6])
def test_dense_tpdm_matches_oracle_c1(n):
    rng = np.random.default_rng(n)
    P = OracleProblem(n, n // 2, n - n // 2, [0] * n, 1)
    S = Structure(n, n // 2, n - n // 2)
    x = rng.standard_normal(S.n_primal)
    assert np.array_equal(S.tpdm(x).reshape(n * n, n * n), P.tpdm(x))
    assert np.array_equal(S.opdm(x, "SUM"), P.opdm(x))


def test_invalid_type_raises():
    S = Structure(4, 2, 2)
    with pytest.raises(Exception):
        S.opdm(np.zeros(S.n_primal), "AB")
