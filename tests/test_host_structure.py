"""Host logic of the product (no GPU): the constraint matrix, b vector,
geminal tables and cost map built by libv2rdm_hip.so agree with the oracle.

Tables and b are integer / exactly representable, so they must agree
bit-exactly.  A u is compared to 1e-13 relative (the product merges
duplicate terms of a row before summing)."""
import ctypes

import numpy as np
import pytest

from hilbert_amd import _native
from hilbert_amd.integrals import synthetic_integrals
from hilbert_amd.structure import Structure
from oracle.pyoracle import OracleProblem

CASES = [
    (4, 2, 2, [0, 0, 0, 0], 1),
    (5, 3, 2, [0] * 5, 1),
    (6, 3, 3, [0, 0, 1, 1, 2, 3], 4),
    (6, 4, 2, [0, 1, 1, 2, 3, 3], 4),
    (7, 3, 3, [0, 1, 2, 3, 4, 5, 7], 8),
    (8, 4, 4, [0] * 8, 1),
]
FLAGS = [(True, True, True, True), (True, True, True, False), (False, True, True, True),
         (True, False, True, True), (True, True, False, True), (False, False, False, False)]


def test_library_exports_every_declared_symbol():
    import re
    import os
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "v2rdm_hip.h")).read()
    declared = set(re.findall(r"\b(v2h_[A-Za-z_0-9]+)\s*\(", hdr))
    declared -= {"v2h_monitor_fn"}
    L = _native.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(_native.SIGNATURES)


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("case", CASES)
def test_constraints_match_oracle(case, flags):
    n, na, nb, irr, H = case
    q2, g2, spin, sz = flags
    S = Structure(n, na, nb, irr, H, q2=q2, g2=g2, spin=spin, sz=sz)
    P = OracleProblem(n, na, nb, irr, H, q2=q2, g2=g2, spin=spin, sz=sz)
    assert (S.n_primal, S.n_dual) == (P.n_primal, P.n_dual)
    assert S.block_dims().tolist() == P.dims.tolist()
    A, b = S.matrix()
    assert np.array_equal(b, P.b())
    rng = np.random.default_rng(1)
    for _ in range(2):
        u = rng.standard_normal(P.n_primal)
        ref = P.Au(u)
        assert np.abs(A @ u - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
        y = rng.standard_normal(P.n_dual)
        ref = P.ATu(y)
        assert np.abs(A.T @ y - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


T1_CASES = CASES[:5]


@pytest.mark.parametrize("t12", [(True, False, False), (False, True, False), (True, True, False), (False, False, True),
                                 (True, True, True)], ids=["T1", "T2", "T1T2", "D3", "T1T2D3"])
@pytest.mark.parametrize("spin_sz", [(True, True), (False, True), (True, False)])
@pytest.mark.parametrize("case", T1_CASES)
def test_t1_t2_d3_constraints_match_oracle(case, spin_sz, t12):
    """POSITIVITY DQGT1 / DQGT2 / DQGT1T2 and CONSTRAIN_D3: T1 blocks after
    G2, T2 blocks after them, D3 blocks last, and the rows of t1.cc, t2.cc's
    slow path and d3.cc."""
    n, na, nb, irr, H = case
    spin, sz = spin_sz
    t1, t2, d3 = t12
    S = Structure(n, na, nb, irr, H, spin=spin, sz=sz, t1=t1, t2=t2, d3=d3)
    P = OracleProblem(n, na, nb, irr, H, spin=spin, sz=sz, t1=t1, t2=t2, d3=d3)
    assert (S.n_primal, S.n_dual) == (P.n_primal, P.n_dual)
    assert S.block_dims().tolist() == P.dims.tolist()
    A, b = S.matrix()
    assert np.array_equal(b, P.b())
    rng = np.random.default_rng(2)
    u = rng.standard_normal(P.n_primal)
    ref = P.Au(u)
    assert np.abs(A @ u - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    y = rng.standard_normal(P.n_dual)
    ref = P.ATu(y)
    assert np.abs(A.T @ y - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_t1_t2_need_q2_and_g2():
    with pytest.raises(Exception, match="Q2 and G2"):
        Structure(4, 2, 2, q2=False, t1=True)
    with pytest.raises(Exception, match="Q2 and G2"):
        Structure(4, 2, 2, g2=False, t2=True)


@pytest.mark.parametrize("case", CASES)
def test_tables_match_oracle(case):
    n, na, nb, irr, H = case
    S = Structure(n, na, nb, irr, H)
    P = OracleProblem(n, na, nb, irr, H)
    ts, to = S.tables(), P.tables()
    for k in ("gems_ab", "gems_aa", "bas_ab", "bas_aa"):
        assert np.array_equal(ts[k], to[k]), k


@pytest.mark.parametrize("case", CASES)
def test_cost_map_matches_oracle(case):
    n, na, nb, irr, H = case
    S = Structure(n, na, nb, irr, H)
    P = OracleProblem(n, na, nb, irr, H)
    h, g = synthetic_integrals(n, 7, 3, irreps=irr if H > 1 else None)
    src = np.concatenate([h.ravel(), g.ravel()])
    i1, i2 = S.cost_map()
    c = np.where(i1 >= 0, src[np.maximum(i1, 0)], 0.0) - np.where(i2 >= 0, src[np.maximum(i2, 0)], 0.0)
    assert np.array_equal(c, P.pack_c(h, g))


def test_orbital_order_is_folded_into_the_cost_map():
    """Irreps given in a non-Pitzer order: the cost vector equals the oracle's
    on the Pitzer-permuted integrals."""
    irr = [2, 0, 1, 0, 3, 1]
    n = len(irr)
    S = Structure(n, 3, 3, irr, 4)
    perm = np.argsort(irr, kind="stable")
    h, g = synthetic_integrals(n, 7, 8, irreps=irr)
    src = np.concatenate([h.ravel(), g.ravel()])
    i1, i2 = S.cost_map()
    c = np.where(i1 >= 0, src[np.maximum(i1, 0)], 0.0) - np.where(i2 >= 0, src[np.maximum(i2, 0)], 0.0)
    P = OracleProblem(n, 3, 3, sorted(irr), 4)
    hp = h[np.ix_(perm, perm)]
    gp = g[np.ix_(perm, perm, perm, perm)]
    assert np.array_equal(c, P.pack_c(hp, gp))


def test_gather_tables_cover_all_entries():
    S = Structure(8, 4, 4)
    st = S.gather_stats()
    # every nonzero appears once in the short or long part of each direction
    A, _ = S.matrix()
    row_len = np.diff(A.indptr)
    short = row_len[row_len <= 32].sum()
    assert st["a_long_entries"] == row_len[row_len > 32].sum()
    assert st["a_entries"] >= short
    col_len = np.diff(A.tocsc().indptr)
    assert st["at_long_entries"] == col_len[col_len > 32].sum()


def test_invalid_descriptions_raise():
    with pytest.raises(_native.V2hError):
        Structure(4, 5, 2)
    with pytest.raises(_native.V2hError):
        Structure(4, 2, 2, [0, 1, 2, 3], 3)
    with pytest.raises(_native.V2hError):
        Structure(4, 2, 2, [0, 1, 5, 3], 4)
