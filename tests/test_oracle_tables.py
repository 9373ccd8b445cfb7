"""Pin the oracle's index tables and vector sizes.

Golden values held by the reference: the archived logs print
"Total number of variables" (n_primal) for three DQG runs
(tests/v2rdm2/output.ref:359, tests/v2rdm4/output.ref:452,3221).  Their
"Total number of constraints" lines come from an older revision with fewer
spin rows; the gap is checked against the row families that revision lacked.
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleProblem


def pitzer(active):
    out = []
    for h, m in enumerate(active):
        out += [h] * m
    return out


LOGGED = [
    # name, active per irrep (D2h), nalpha, nbeta, logged n_primal, logged n_dual
    ("v2rdm2", [1, 0, 1, 1, 0, 1, 1, 1], 3, 3, 1668, 1567),
    ("v2rdm4 singlet", [0, 1, 1, 3, 1, 3, 0, 1], 5, 5, 15508, 14355),
    ("v2rdm4 triplet", [0, 1, 1, 3, 1, 3, 0, 1], 6, 4, 20236, 18134),
]


@pytest.mark.parametrize("name,active,na,nb,np_log,nd_log", LOGGED)
def test_n_primal_matches_reference_logs(name, active, na, nb, np_log, nd_log):
    P = OracleProblem(sum(active), na, nb, pitzer(active), 8)
    assert P.n_primal == np_log
    gab, gaa = P.gems()
    sg = int((gab.astype(np.int64) ** 2).sum())
    sa = int((gaa.astype(np.int64) ** 2).sum())
    # the older revision lacked the D2aa/D2bb <- D2ab and D200/D2ab-symmetry
    # singlet rows (2 sg + 2 sa), or sg + 2 sa rows in the non-singlet case
    missing = 2 * sg + 2 * sa if na == nb else sg + 2 * sa
    assert P.n_dual - missing == nd_log


@pytest.mark.parametrize("n,na,irr,H,npr,ndu", [
    (8, 4, [0] * 8, 1, 40256, 47028),           # Hubbard 8-site
    (10, 5, [0] * 10, 1, 98500, 115079),        # (10e,10o) C1
    (16, 8, sorted([0, 1, 2, 3] * 4), 4, 162304, 189956),  # (16e,16o) C2v
])
def test_config_sizes(n, na, irr, H, npr, ndu):
    P = OracleProblem(n, na, na, irr, H)
    assert (P.n_primal, P.n_dual) == (npr, ndu)


def _random_irreps(rng, n, H):
    return sorted(rng.integers(0, H, n).tolist())


@pytest.mark.parametrize("seed", range(6))
def test_table_invariants(seed):
    rng = np.random.default_rng(seed)
    H = [1, 2, 4, 8][seed % 4]
    n = int(rng.integers(3, 9))
    irr = _random_irreps(rng, n, H)
    P = OracleProblem(n, n // 2, n - n // 2, irr, H)
    t = P.tables()
    m = np.bincount(irr, minlength=H)
    start_ab = np.concatenate([[0], np.cumsum(t["gems_ab"])[:-1]])
    start_aa = np.concatenate([[0], np.cumsum(t["gems_aa"])[:-1]])
    for h in range(H):
        # geminal counts from the direct-product rule
        assert t["gems_ab"][h] == sum(m[a] * m[a ^ h] for a in range(H))
        assert t["gems_aa"][h] == (t["gems_ab"][h] - (n if h == 0 else 0)) // 2
        pairs = t["bas_ab"][start_ab[h]:start_ab[h] + t["gems_ab"][h]]
        for k, (i, j) in enumerate(pairs):
            assert irr[i] ^ irr[j] == h
            assert t["ibas_ab"][h, i, j] == k
        # pair order: by second orbital, then first
        keys = [(j, i) for i, j in pairs]
        assert keys == sorted(keys)
        apairs = t["bas_aa"][start_aa[h]:start_aa[h] + t["gems_aa"][h]]
        for k, (i, j) in enumerate(apairs):
            assert i < j
            assert t["ibas_aa"][h, i, j] == k and t["ibas_aa"][h, j, i] == k
        keys = [(j, i) for i, j in apairs]
        assert keys == sorted(keys)
        # every pair outside irrep h is marked -999
        outside = (np.asarray(irr)[:, None] ^ np.asarray(irr)[None, :]) != h
        assert (t["ibas_ab"][h][outside] == -999).all()


def test_c1_closed_form():
    n = 7
    P = OracleProblem(n, 3, 3)
    t = P.tables()
    for mm in range(n * n):
        assert tuple(t["bas_ab"][mm]) == (mm % n, mm // n)
    expect = [(p, q) for q in range(n) for p in range(q)]
    assert [tuple(v) for v in t["bas_aa"]] == expect


def test_block_dims_and_offsets_are_contiguous():
    P = OracleProblem(6, 3, 2, [0, 0, 1, 1, 2, 3], 4)
    off = P.offsets()
    gab, gaa = P.gems()
    # blocks in layout order: D2ab, D2aa, D2bb, D200 (2g, non-singlet), D1a, D1b,
    # Q1a, Q1b, Q2ab, Q2aa, Q2bb, G2ab, G2ba, G2aa (2g)
    m = np.bincount([0, 0, 1, 1, 2, 3], minlength=4)
    expect = list(gab) + list(gaa) + list(gaa) + list(2 * gab) + list(m) * 4 \
        + list(gab) + list(gaa) + list(gaa) + list(gab) + list(gab) + list(2 * gab)
    assert list(P.dims) == expect
    assert sum(int(d) ** 2 for d in P.dims) == P.n_primal
    assert off["d2ab"][0] == 0
    assert off["g2aa"][3] + (2 * gab[3]) ** 2 == P.n_primal
