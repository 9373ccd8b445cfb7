"""Pin the oracle's density-fitted intake (orc_df_integrals) without the reference.

The oracle restates the DF branch of TEI, FrozenCoreEnergy and the integrals
RepackIntegrals reads (tei.cc:139-283).  What fixes it here:

* energy of an embedded state: for a random full-space Hamiltonian given as
  DF integrals, and any state Phi = (doubly occupied core) x Psi_act with
  Psi_act a random spin-adapted active-space state, the brute-force
  full-space energy <Phi|H|Phi> (tests/fci.py, with (pq|rs) formed from the
  unpacked three-index tensor) equals E_fzc + c(h', g') . x(Psi_act).  This
  checks the packed-triangle addressing, the factor 2 and the sign of the
  Coulomb and exchange core terms and E_fzc together; core and active
  orbitals are scattered over the full space and the active list is not
  sorted, so a wrong index map fails it;
* the closed-shell core determinant alone has energy E_fzc;
* with no core, g' is the plain DF product and h' the active block of h.
"""
import numpy as np
import pytest

from fci import Fock, RDMs, pack_primal, spin_adapted_random_state
from hilbert_amd.integrals import synthetic_df
from oracle.pyoracle import OracleProblem, df_integrals

_FOCK = {}


def _fock(n):
    if n not in _FOCK:
        _FOCK[n] = Fock(n)
    return _FOCK[n]


def _full_hamiltonian(h, L, perm):
    """Brute-force H over the orbitals relabelled by perm (new k = old perm[k])."""
    g = np.einsum("Qpq,Qrs->pqrs", L, L)
    hp = h[np.ix_(perm, perm)]
    gp = g[np.ix_(perm, perm, perm, perm)]
    return _fock(len(perm)).hamiltonian(hp, gp)


def _embed(psi_act, nact, ncore, nfull):
    """Core orbitals are modes 0..ncore-1 of each spin, active ones follow in
    order: the map keeps the mode order, so the embedded vector is
    +- (core creators) Psi_act for any determinant sign convention."""
    out = np.zeros(1 << (2 * nfull))
    core_bits = 0
    for k in range(ncore):
        core_bits |= 1 << k
        core_bits |= 1 << (nfull + k)
    for det in np.nonzero(psi_act)[0]:
        full = core_bits
        for p in range(nact):
            if det >> p & 1:
                full |= 1 << (ncore + p)
            if det >> (nact + p) & 1:
                full |= 1 << (nfull + ncore + p)
        out[full] = psi_act[det]
    return out


CASES = [
    # nfull, core, active (full-space indices, active order), na, nb, naux
    (5, [3], [4, 0, 2], 2, 1, 7),
    (5, [1, 4], [2, 0], 1, 1, 6),
    (5, [0, 2], [4, 3, 1], 2, 2, 9),
    (4, [], [3, 1, 0, 2], 2, 2, 5),
]


@pytest.mark.parametrize("case", CASES)
def test_embedded_state_energy(case):
    nfull, core, active, na, nb, naux = case
    rng = np.random.default_rng(31 + nfull + len(core))
    h, qmo, L = synthetic_df(nfull, naux, rng)
    virt = [k for k in range(nfull) if k not in core and k not in active]
    perm = list(core) + list(active) + virt
    H = _full_hamiltonian(h, L, perm)
    nact = len(active)
    Fa = _fock(nact)
    psi = spin_adapted_random_state(Fa, na, nb, rng)
    phi = _embed(psi, nact, len(core), nfull)
    e_full = phi @ (H @ phi)
    ha, ga, efzc = df_integrals(h, qmo, core, active)
    P = OracleProblem(nact, na, nb)
    e_split = efzc + P.pack_c(ha, ga) @ pack_primal(P, RDMs(Fa, psi))
    assert abs(e_split - e_full) < 1e-11 * max(1.0, abs(e_full))


def test_core_determinant_energy():
    nfull, core = 5, [1, 3, 4]
    rng = np.random.default_rng(8)
    h, qmo, L = synthetic_df(nfull, 8, rng)
    perm = core + [0, 2]
    H = _full_hamiltonian(h, L, perm)
    det = 0
    for k in range(len(core)):
        det |= 1 << k
        det |= 1 << (nfull + k)
    _, _, efzc = df_integrals(h, qmo, core, [0])
    assert abs(efzc - H[det, det]) < 1e-12 * max(1.0, abs(efzc))


def test_no_core_is_the_df_product():
    rng = np.random.default_rng(9)
    h, qmo, L = synthetic_df(6, 11, rng)
    active = [5, 1, 3, 0]
    ha, ga, efzc = df_integrals(h, qmo, [], active)
    g = np.einsum("Qpq,Qrs->pqrs", L, L)
    assert efzc == 0.0
    assert np.array_equal(ha, h[np.ix_(active, active)])
    assert np.abs(ga - g[np.ix_(active, active, active, active)]).max() < 1e-14
