"""The benchmark / parity workloads, generated from fixed seeds.

hubbard8      8-site periodic Hubbard ring, 8 electrons, t = 1, U = 4, C1
c1_10         (10e,10o) C1, synthetic integrals, naux = 30
batch64_d2h   64 SDPs, n ~ U{6..12}, electrons = n (odd n: nalpha = nbeta + 1),
              D2h labels at random, naux = 30
c2v_16        (16e,16o), 4 irreps (C2v labels at random), naux = 48
batch16_big   16 SDPs, n ~ U{20..24}, D2h labels at random, naux = 3 n

Every instance is (Problem, h, g): symmetry-adapted integrals in the
orbital order of the labels (which are returned sorted, i.e. Pitzer order).
"""
from __future__ import annotations

import numpy as np

from .batch import Problem
from .integrals import hubbard_integrals, random_irreps, synthetic_integrals

SEED = 20261202


def _electrons(n):
    return (n + 1) // 2, n // 2


def hubbard8():
    h, g = hubbard_integrals(8, 1.0, 4.0)
    return [(Problem(8, 4, 4), h, g)]


def synthetic(n, naux, nirrep, seed, nalpha=None, nbeta=None):
    rng = np.random.default_rng(seed)
    irr = random_irreps(n, nirrep, rng)
    h, g = synthetic_integrals(n, naux, rng, irreps=irr if nirrep > 1 else None)
    na, nb = _electrons(n) if nalpha is None else (nalpha, nbeta)
    return Problem(n, na, nb, irreps=irr.tolist() if nirrep > 1 else None, nirrep=nirrep), h, g


def c1_10(seed=SEED + 1):
    return [synthetic(10, 30, 1, seed)]


def c2v(n, count=1, seed=SEED + 3):
    """(ne,no) with n electrons in n orbitals, C2v labels at random, naux 3 n."""
    return [synthetic(n, 3 * n, 4, seed + k) for k in range(count)]


def c2v_16(count=1, seed=SEED + 3):
    return c2v(16, count, seed)


def batch64_d2h(count=64, seed=SEED + 2):
    rng = np.random.default_rng(seed)
    ns = rng.integers(6, 13, count)
    return [synthetic(int(n), 30, 8, seed * 1000 + k) for k, n in enumerate(ns)]


def batch16_big(count=16, seed=SEED + 4):
    rng = np.random.default_rng(seed)
    ns = rng.integers(20, 25, count)
    return [synthetic(int(n), 3 * int(n), 8, seed * 1000 + k) for k, n in enumerate(ns)]


CONFIGS = {"hubbard8": hubbard8, "c1_10": c1_10, "batch64_d2h": batch64_d2h, "c2v_16": c2v_16,
           "batch16_big": batch16_big}
