"""Pin the oracle's constraint maps and c-packing without the reference.

* adjointness: <y, A u> = <A^T y, u> (checks the separately written Au and
  ATu loops against each other);
* exact-RDM feasibility: for brute-force N-electron states with S = Ms and
  definite spatial symmetry, every DQG+spin row holds, A x(Psi) = b, with
  x(Psi) built from second-quantised expectation values (tests/fci.py);
* c.x(Psi) = <Psi|H|Psi> for symmetry-adapted random integrals;
* pack_c_adjoint is the exact transpose of pack_c.
"""
import itertools

import numpy as np
import pytest

from fci import Fock, RDMs, pack_primal, spin_adapted_random_state
from hilbert_amd.integrals import synthetic_integrals
from oracle.pyoracle import OracleProblem

CASES = [
    # n, na, nb, irreps, nirrep
    (4, 2, 2, [0, 0, 0, 0], 1),
    (4, 2, 1, [0, 0, 0, 0], 1),
    (4, 3, 1, [0, 0, 0, 0], 1),
    (5, 2, 2, [0, 0, 1, 1, 1], 2),
    (4, 2, 2, [0, 1, 2, 3], 4),
    (5, 3, 2, [0, 1, 1, 2, 3], 4),
]
FLAGS = list(itertools.product([True, False], repeat=4))  # q2, g2, spin, sz


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("case", CASES)
def test_adjointness(case, flags):
    n, na, nb, irr, H = case
    P = OracleProblem(n, na, nb, irr, H, q2=flags[0], g2=flags[1], spin=flags[2], sz=flags[3])
    rng = np.random.default_rng(7)
    u = rng.standard_normal(P.n_primal)
    y = rng.standard_normal(P.n_dual)
    lhs = y @ P.Au(u)
    rhs = P.ATu(y) @ u
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


_FOCK = {}


def _fock(n):
    if n not in _FOCK:
        _FOCK[n] = Fock(n)
    return _FOCK[n]


@pytest.mark.parametrize("flags", [(True, True, True, True), (True, True, True, False),
                                   (False, True, True, True), (True, False, False, True)])
@pytest.mark.parametrize("case", CASES)
def test_exact_rdm_feasibility(case, flags):
    n, na, nb, irr, H = case
    F = _fock(n)
    rng = np.random.default_rng(11)
    psi = spin_adapted_random_state(F, na, nb, rng, irreps=irr if H > 1 else None, gamma=0)
    P = OracleProblem(n, na, nb, irr, H, q2=flags[0], g2=flags[1], spin=flags[2], sz=flags[3])
    x = pack_primal(P, RDMs(F, psi))
    res = P.Au(x) - P.b()
    assert np.abs(res).max() < 1e-11


def test_feasibility_detects_wrong_spin_state():
    """A triplet state in the na = nb sector violates the singlet spin rows."""
    n, na, nb = 4, 2, 2
    F = _fock(n)
    sec = F.sector(na, nb)
    S2 = F.s_squared()[sec][:, sec].toarray()
    w, v = np.linalg.eigh(S2)
    psi = np.zeros(F.dim)
    psi[sec] = v[:, np.argmin(np.abs(w - 2.0))]
    P = OracleProblem(n, na, nb)
    res = P.Au(pack_primal(P, RDMs(F, psi))) - P.b()
    assert np.abs(res).max() > 1e-3


@pytest.mark.parametrize("case", CASES)
def test_energy_functional(case):
    n, na, nb, irr, H = case
    F = _fock(n)
    rng = np.random.default_rng(5)
    h, g = synthetic_integrals(n, 6, rng, irreps=irr if H > 1 else None)
    psi = spin_adapted_random_state(F, na, nb, rng, irreps=irr if H > 1 else None, gamma=0)
    e_exact = psi @ (F.hamiltonian(h, g) @ psi)
    P = OracleProblem(n, na, nb, irr, H)
    x = pack_primal(P, RDMs(F, psi))
    assert abs(P.pack_c(h, g) @ x - e_exact) < 1e-11 * max(1.0, abs(e_exact))


@pytest.mark.parametrize("case", CASES[:4])
def test_pack_c_adjoint_is_transpose(case):
    n, na, nb, irr, H = case
    P = OracleProblem(n, na, nb, irr, H)
    rng = np.random.default_rng(2)
    h = rng.standard_normal((n, n))
    g = rng.standard_normal((n,) * 4)
    xb = rng.standard_normal(P.n_primal)
    hb, gb = P.pack_c_adjoint(xb)
    lhs = P.pack_c(h, g) @ xb
    rhs = (h * hb).sum() + (g * gb).sum()
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_b_vector_values():
    n, na, nb = 4, 2, 1
    P = OracleProblem(n, na, nb)
    b = P.b()
    assert b[:3].tolist() == [na * nb, na * (na - 1), nb * (nb - 1)]
    ms = 0.5
    spin = 0.5 * (na + nb) + ms * ms - ms * (ms + 1)
    # D1+Q1 = 1 on 2n diagonal rows; Q2ab = -1 on n^2 diagonal geminals;
    # Q2aa, Q2bb = -1 on n(n-1)/2 diagonal geminals each
    expect = na * nb + na * (na - 1) + nb * (nb - 1) + 2 * n + spin - n * n - n * (n - 1)
    assert np.isclose(b.sum(), expect)
    assert np.count_nonzero(b) == 3 - (nb * (nb - 1) == 0) + 2 * n + 1 + n * n + n * (n - 1)
