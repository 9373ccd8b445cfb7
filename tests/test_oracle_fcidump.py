"""Pin the oracle's FCIDUMP records (orc_fcidump_*) without the reference.

The records restate fcidump.cc:60-243: unique (pq|rs) from the DF integrals,
the one-electron integrals by irrep block, the nuclear repulsion, then the
spin-free 2-RDM and the spin-summed 1-RDM of the active space with labels
ordered core, active, virtual.  What fixes them here:

* read back like a consumer would (8-fold expansion of the integrals, labels
  1..ninact as the doubly occupied core, 2-RDM records e2 += dum (ik|jl) / 2
  as fcidump.cc:212 evaluates them), the files give the brute-force energy
  <Phi|H|Phi> + enuc of an embedded state Phi = core x Psi_act over the full
  space, with core, active and virtual orbitals scattered in the full order;
* the record counts match the closed forms, with and without symmetry, and
  every integral record has equal pair symmetry and canonical order;
* labels: core first, then active, then virtual, each by irrep then index.
"""
import numpy as np
import pytest

from fci import Fock, RDMs, pack_primal, spin_adapted_random_state
from hilbert_amd.integrals import synthetic_df
from oracle.pyoracle import OracleProblem, fcidump_integrals, fcidump_map

_FOCK = {}


def _fock(n):
    if n not in _FOCK:
        _FOCK[n] = Fock(n)
    return _FOCK[n]


def _embed(psi_act, nact, ncore, nfull):
    out = np.zeros(1 << (2 * nfull))
    core_bits = 0
    for k in range(ncore):
        core_bits |= (1 << k) | (1 << (nfull + k))
    for det in np.nonzero(psi_act)[0]:
        full = core_bits
        for p in range(nact):
            if det >> p & 1:
                full |= 1 << (ncore + p)
            if det >> (nact + p) & 1:
                full |= 1 << (nfull + ncore + p)
        out[full] = psi_act[det]
    return out


def read_back(ival, ilab, rval, rlab, nfull, ninact, nact):
    """Energy from the records, as a consumer of the two files computes it."""
    enuc = None
    h = np.zeros((nfull, nfull))
    g = np.zeros((nfull,) * 4)
    for v, (p, q, r, s) in zip(ival, ilab):
        if p == 0:
            enuc = v
        elif r == 0:
            h[p - 1, q - 1] = v
        else:
            p, q, r, s = p - 1, q - 1, r - 1, s - 1
            for a, b, c, d in ((p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r)):
                g[a, b, c, d] = v
                g[c, d, a, b] = v
    core = list(range(ninact))
    efzc = sum(2 * h[c, c] for c in core) + sum(2 * g[c, c, d, d] - g[c, d, c, d] for c in core for d in core)
    e = enuc + efzc
    for v, (i, j, k, l) in zip(rval, rlab):
        i, j = i - 1 + ninact, j - 1 + ninact
        if k == 0:
            hp = h[i, j] + sum(2 * g[i, j, c, c] - g[i, c, j, c] for c in core)
            e += v * hp
        else:
            k, l = k - 1 + ninact, l - 1 + ninact
            e += 0.5 * v * g[i, k, j, l]
    return e


CASES = [
    # nfull, core, active (Pitzer active order), na, nb, naux
    (5, [3], [0, 2, 4], 2, 1, 7),
    (5, [1, 4], [0, 2], 1, 1, 6),
    (4, [], [1, 2, 3], 2, 1, 5),
]


@pytest.mark.parametrize("case", CASES)
def test_records_give_the_brute_force_energy(case):
    nfull, core, active, na, nb, naux = case
    rng = np.random.default_rng(77 + nfull + len(core))
    h, qmo, L = synthetic_df(nfull, naux, rng)
    enuc = 3.25
    space = np.full(nfull, 2, dtype=np.int32)
    space[core] = 0
    space[active] = 1
    lab = fcidump_map(space)
    ival, ilab = fcidump_integrals(h, qmo, enuc, lab)
    nact = len(active)
    P = OracleProblem(nact, na, nb)
    Fa = _fock(nact)
    psi = spin_adapted_random_state(Fa, na, nb, rng)
    x = pack_primal(P, RDMs(Fa, psi))
    rval, rlab = P.fcidump_rdms(x, active, lab, len(core))
    e_files = read_back(ival, ilab, rval, rlab, nfull, len(core), nact)
    # brute force over the full space, orbitals relabelled core, active, virtual
    virt = [k for k in range(nfull) if k not in core and k not in active]
    perm = list(core) + list(active) + virt
    g = np.einsum("Qpq,Qrs->pqrs", L, L)
    H = _fock(nfull).hamiltonian(h[np.ix_(perm, perm)], g[np.ix_(perm, perm, perm, perm)])
    phi = _embed(psi, nact, len(core), nfull)
    assert abs(e_files - (phi @ (H @ phi) + enuc)) < 1e-11 * max(1.0, abs(e_files))


def test_counts_symmetry_and_labels():
    nfull, nirrep = 7, 4
    irr = np.array([0, 0, 1, 2, 2, 3, 3], dtype=np.int32)
    space = np.array([0, 1, 1, 2, 1, 0, 1], dtype=np.int32)
    rng = np.random.default_rng(5)
    h, qmo, _ = synthetic_df(nfull, 9, rng, irreps=irr)
    lab = fcidump_map(space, irr, nirrep)
    # core by irrep (orbitals 0, 5), then active by irrep (1, 2, 4, 6), then virtual (3)
    assert lab.tolist() == [1, 3, 4, 7, 5, 2, 6]
    val, labs = fcidump_integrals(h, qmo, 1.5, lab, irr)
    inv = np.argsort(lab)                      # label - 1 -> orbital
    pairs = [(p, q) for p in range(nfull) for q in range(p, nfull)]
    expect = sum(1 for a in pairs for b in pairs
                 if (irr[a[0]] ^ irr[a[1]]) == (irr[b[0]] ^ irr[b[1]])
                 and (a[1] * (a[1] + 1) // 2 + a[0]) <= (b[1] * (b[1] + 1) // 2 + b[0]))
    one = sum(int((irr == k).sum()) ** 2 for k in range(nirrep))
    assert len(val) == expect + one + 1
    two = labs[:expect] - 1
    o = inv[two]
    assert np.all((irr[o[:, 0]] ^ irr[o[:, 1]]) == (irr[o[:, 2]] ^ irr[o[:, 3]]))
    assert np.all(o[:, 0] <= o[:, 1]) and np.all(o[:, 2] <= o[:, 3])
    assert np.all(labs[expect:expect + one, 2:] == 0) and labs[-1].tolist() == [0, 0, 0, 0] and val[-1] == 1.5


def test_rdm_record_count_and_labels():
    P = OracleProblem(4, 2, 2, [0, 0, 1, 1], 2)
    t = P.tables()
    x = np.arange(P.n_primal, dtype=float)
    act = np.array([2, 3, 5, 6], dtype=np.int32)
    space = np.array([0, 0, 1, 1, 2, 1, 1], dtype=np.int32)
    lab = fcidump_map(space)
    val, labs = P.fcidump_rdms(x, act, lab, 2)
    gab = t["gems_ab"]
    n2 = int((gab.astype(np.int64) ** 2).sum())
    assert len(val) == n2 + int(gab[0])
    assert labs.min() >= 0 and labs[:, :2].min() >= 1 and labs.max() <= 4
