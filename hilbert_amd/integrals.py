"""Hamiltonians for the fixed-orbital v2RDM solve.

* ``hubbard_integrals`` -- the periodic 1D Hubbard ring hilbert builds from
  HUBBARD_T / HUBBARD_U (oei.cc:76-95, tei.cc:285-291).
* ``synthetic_integrals`` -- molecular-like integrals from a seed:
  h symmetric with diagonal -U(0.5, 2.5) and off-diagonal N(0, 0.1);
  (pq|rs) = sum_Q L^Q_pq L^Q_rs with L^Q symmetric, entries N(0, 0.15),
  so g has 8-fold symmetry and is positive semidefinite as a pq x rs matrix.
  With orbital irrep labels the integrals are symmetry adapted: h_pq = 0
  unless sym(p) = sym(q), every auxiliary function Q carries an irrep and
  L^Q_pq = 0 unless sym(p) x sym(q) = sym(Q).

* ``synthetic_df`` / ``pack_qmo`` -- the same construction over a full
  orbital space, returned as density-fitted three-index integrals in the
  packed (Q|mn) layout of the reference's Qmo_ (tei.cc:280-283): row Q holds
  L^Q_mn at INDEX(m, n) = m (m + 1) / 2 + n for m >= n.

Integral layout everywhere: h[p, q], g[p, q, r, s] = (pq|rs) (chemists'
notation), the layout of hilbert's external-Hamiltonian constructor
(v2rdm_solver.cc:147-195).
"""
from __future__ import annotations

import numpy as np


def hubbard_integrals(n_sites: int, t: float, U: float):
    h = np.zeros((n_sites, n_sites))
    for i in range(n_sites):
        for j in range(n_sites):
            d = abs(i - j)
            if d == 1 or d == n_sites - 1:
                h[i, j] = -t
    g = np.zeros((n_sites,) * 4)
    for i in range(n_sites):
        g[i, i, i, i] = U
    return h, g


def random_irreps(n: int, nirrep: int, rng) -> np.ndarray:
    """Uniformly random irrep labels, returned in Pitzer (sorted) order."""
    if nirrep == 1:
        return np.zeros(n, dtype=np.int32)
    return np.sort(rng.integers(0, nirrep, n)).astype(np.int32)


def synthetic_integrals(n: int, naux: int, rng, irreps=None):
    rng = np.random.default_rng(rng)
    h = rng.normal(0.0, 0.1, (n, n))
    h = np.triu(h, 1)
    h = h + h.T
    h[np.diag_indices(n)] = -rng.uniform(0.5, 2.5, n)
    L = rng.normal(0.0, 0.15, (naux, n, n))
    L = np.triu(L) + np.transpose(np.triu(L, 1), (0, 2, 1))
    if irreps is not None:
        irr = np.asarray(irreps)
        nirrep = 1
        while nirrep <= int(irr.max()):
            nirrep *= 2
        h[irr[:, None] != irr[None, :]] = 0.0
        qirr = rng.integers(0, nirrep, naux)
        pair = irr[:, None] ^ irr[None, :]
        for Q in range(naux):
            L[Q][pair != qirr[Q]] = 0.0
    Lm = L.reshape(naux, n * n)
    g = (Lm.T @ Lm).reshape(n, n, n, n)
    return h, g


def pack_qmo(L):
    """(nq, n, n) symmetric three-index integrals -> (nq, n (n + 1) / 2),
    packed lower triangle, row-major: (0,0), (1,0), (1,1), (2,0), ..."""
    L = np.asarray(L, dtype=np.float64)
    r, c = np.tril_indices(L.shape[1])
    return np.ascontiguousarray(L[:, r, c])


def synthetic_df(nfull: int, naux: int, rng, irreps=None):
    """(h_full, qmo, L): synthetic_integrals' h and L over nfull orbitals, with
    L also packed as qmo."""
    rng = np.random.default_rng(rng)
    h = rng.normal(0.0, 0.1, (nfull, nfull))
    h = np.triu(h, 1)
    h = h + h.T
    h[np.diag_indices(nfull)] = -rng.uniform(0.5, 2.5, nfull)
    L = rng.normal(0.0, 0.15, (naux, nfull, nfull))
    L = np.triu(L) + np.transpose(np.triu(L, 1), (0, 2, 1))
    if irreps is not None:
        irr = np.asarray(irreps)
        nirrep = 1
        while nirrep <= int(irr.max()):
            nirrep *= 2
        h[irr[:, None] != irr[None, :]] = 0.0
        qirr = rng.integers(0, nirrep, naux)
        pair = irr[:, None] ^ irr[None, :]
        for Q in range(naux):
            L[Q][pair != qirr[Q]] = 0.0
    return h, pack_qmo(L), L
