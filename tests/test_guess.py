"""TPDM_GUESS HF (Guess, v2rdm_solver.cc:1711-1798; q2.cc:45-167; g2.cc:44-242),
host side: for a closed-shell determinant the guess is exactly the packed
RDMs of that determinant (D1, Q1, D2, Q2, G2 from brute-force second
quantisation), with the D200 blocks left zero as in the reference."""
import numpy as np
import pytest

from fci import Fock, RDMs, pack_primal, pack_t1, pack_t2
from hilbert_amd.structure import Structure
from oracle.pyoracle import OracleProblem

CASES = [
    (4, 2, None, 1),
    (5, 2, None, 1),
    (6, 3, [0, 0, 1, 1, 2, 3], 4),
    (6, 2, [0, 1, 2, 3, 5, 6], 8),
]


def _determinant(F, n, na, nb):
    idx = sum(1 << p for p in range(na)) + sum(1 << (p + n) for p in range(nb))
    psi = np.zeros(F.dim)
    psi[idx] = 1.0
    return psi


@pytest.mark.parametrize("n,na,irr,H", CASES)
def test_hf_guess_is_the_determinant(n, na, irr, H):
    irreps = irr if irr is not None else [0] * n
    P = OracleProblem(n, na, na, irreps, H)
    S = Structure(n, na, na, irr, H)
    F = Fock(n)
    x_ref = pack_primal(P, RDMs(F, _determinant(F, n, na, na)))
    off = P.offsets()
    dims = S.block_dims()
    # zero the D200 blocks (families: D2ab, D2aa, D2bb, D200, ... per irrep)
    for h in range(H):
        d = dims[3 * H + h]
        x_ref[off["d200"][h]:off["d200"][h] + d * d] = 0.0
    x = S.hf_guess()
    assert np.abs(x - x_ref).max() < 1e-14


def test_hf_guess_open_shell_follows_reference_occupations():
    """The external path sets doccpi = nalpha (v2rdm_solver.cc:1004-1005), so
    the guess occupies nalpha beta orbitals as well."""
    n, na, nb = 5, 3, 2
    S = Structure(n, na, nb)
    P = OracleProblem(n, na, nb, [0] * n, 1)
    x = S.hf_guess()
    off = P.offsets()
    d1b = x[off["d1b"][0]:off["d1b"][0] + n * n].reshape(n, n)
    d1a = x[off["d1a"][0]:off["d1a"][0] + n * n].reshape(n, n)
    assert np.trace(d1a) == na and np.trace(d1b) == na


@pytest.mark.parametrize("n,na,irr,H", CASES[:3])
def test_hf_guess_with_t1_is_the_determinant(n, na, irr, H):
    """POSITIVITY DQGT1: the T1 blocks of the guess (each T1 row solved for
    its T1 element once the Q2 and G2 rows have fixed theirs) are the exact
    T1 of the determinant."""
    irreps = irr if irr is not None else [0] * n
    P = OracleProblem(n, na, na, irreps, H, t1=True)
    P0 = OracleProblem(n, na, na, irreps, H)
    S = Structure(n, na, na, irr, H, t1=True)
    F = Fock(n)
    psi = _determinant(F, n, na, na)
    x_ref = np.zeros(P.n_primal)
    x_ref[:P0.n_primal] = pack_primal(P0, RDMs(F, psi))
    pack_t1(P, F, psi, x_ref)
    off = P.offsets()
    dims = S.block_dims()
    for h in range(H):
        d = dims[3 * H + h]
        x_ref[off["d200"][h]:off["d200"][h] + d * d] = 0.0
    x = S.hf_guess()
    assert np.abs(x - x_ref).max() < 1e-13


@pytest.mark.parametrize("n,na,irr,H", CASES[:3])
def test_hf_guess_with_t1_t2_is_the_determinant(n, na, irr, H):
    """POSITIVITY DQGT1T2: T1 and T2 blocks of the guess are those of the
    determinant (every T2 row holds one T2 element besides D1 and D2)."""
    irreps = irr if irr is not None else [0] * n
    P = OracleProblem(n, na, na, irreps, H, t1=True, t2=True)
    P0 = OracleProblem(n, na, na, irreps, H)
    S = Structure(n, na, na, irr, H, t1=True, t2=True)
    F = Fock(n)
    psi = _determinant(F, n, na, na)
    x_ref = np.zeros(P.n_primal)
    x_ref[:P0.n_primal] = pack_primal(P0, RDMs(F, psi))
    pack_t1(P, F, psi, x_ref)
    pack_t2(P, F, psi, x_ref)
    off = P.offsets()
    dims = S.block_dims()
    for h in range(H):
        d = dims[3 * H + h]
        x_ref[off["d200"][h]:off["d200"][h] + d * d] = 0.0
    assert np.abs(S.hf_guess() - x_ref).max() < 1e-13
