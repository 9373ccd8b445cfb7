"""GPU parity: the HIP path, called through the C ABI, against the oracle.

* maps:  A u and A^T y to 1e-13 (relative to max |ref|), every structure of a
  mixed batch (C1, 2/4/8 irreps, singlet / doublet / triplet);
* cost packing bit-exact; its adjoint to 1e-13;
* full solves at e_convergence 1e-7 / r_convergence 1e-6 (both sides):
  total energy within 1e-6 Eh of the oracle; at 1e-9 / 1e-8 (where the
  RDMs are determined well enough to compare): energy within 1e-8 Eh, 1-RDM
  and <S^2> within 1e-5;
* full solves at the default 1e-6 / 1e-5: the stopping criteria hold and
  the energy lies inside the band those thresholds allow (5e-5 Eh); CG on the
  singular A A^T is chaotic in finite precision, so two correct
  implementations stop at different points of that band;
* a batch returns bit-identical answers to solving each instance alone.
"""
import json
import os

import numpy as np
import pytest

from hilbert_amd.batch import Batch, Problem, options
from hilbert_amd.integrals import hubbard_integrals, synthetic_integrals
from oracle.pyoracle import OracleProblem

pytestmark = pytest.mark.gpu

MAP_CASES = [
    (4, 2, 2, None, 1),
    (5, 3, 2, None, 1),
    (6, 3, 3, [0, 0, 1, 1, 2, 3], 4),
    (6, 4, 2, [0, 1, 1, 2, 3, 3], 4),
    (7, 3, 3, [0, 1, 2, 3, 4, 5, 7], 8),
    (8, 4, 4, None, 1),
    (10, 5, 5, None, 1),
    (9, 5, 4, [0, 0, 1, 2, 2, 3, 5, 6, 7], 8),
]


def _oracle(n, na, nb, irr, H):
    return OracleProblem(n, na, nb, irr if irr is not None else [0] * n, H)


@pytest.fixture(scope="module")
def map_batch():
    probs = [Problem(n, na, nb, irr, H) for n, na, nb, irr, H in MAP_CASES]
    B = Batch(probs, options())
    yield B
    B.close()


def test_batch_sizes_match_oracle(map_batch):
    for i, case in enumerate(MAP_CASES):
        P = _oracle(*case)
        assert map_batch.n_primal[i] == P.n_primal
        assert map_batch.n_dual[i] == P.n_dual
        assert map_batch.block_dims(i).tolist() == P.dims.tolist()
        assert np.array_equal(map_batch.b(i), P.b())


def test_maps_match_oracle(map_batch):
    rng = np.random.default_rng(3)
    for i, case in enumerate(MAP_CASES):
        P = _oracle(*case)
        for _ in range(2):
            u = rng.standard_normal(P.n_primal)
            ref = P.Au(u)
            got = map_batch.Au(i, u)
            assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), case
            y = rng.standard_normal(P.n_dual)
            ref = P.ATu(y)
            got = map_batch.ATu(i, y)
            assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), case


def test_cost_packing_and_adjoint(map_batch):
    rng = np.random.default_rng(4)
    for i, (n, na, nb, irr, H) in enumerate(MAP_CASES):
        P = _oracle(n, na, nb, irr, H)
        h, g = synthetic_integrals(n, 9, rng, irreps=irr)
        map_batch.set_integrals(i, h, g)
        assert np.array_equal(map_batch.c(i), P.pack_c(h, g))
        xb = rng.standard_normal(P.n_primal)
        dh, dg = map_batch.cost_adjoint(i, xb)
        rh, rg = P.pack_c_adjoint(xb)
        assert np.abs(dh - rh).max() <= 1e-13 * max(1.0, np.abs(rh).max())
        assert np.abs(dg - rg).max() <= 1e-13 * max(1.0, np.abs(rg).max())


SOLVE_CASES = [
    ("hubbard2", lambda: (Problem(2, 1, 1),) + hubbard_integrals(2, 1.0, 4.0)),
    ("c1_n4", lambda: (Problem(4, 2, 2),) + synthetic_integrals(4, 8, 1)),
    ("doublet_n5", lambda: (Problem(5, 3, 2),) + synthetic_integrals(5, 10, 2)),
    ("c2v_n6", lambda: (Problem(6, 3, 3, [0, 0, 1, 2, 3, 3], 4),)
     + synthetic_integrals(6, 12, 3, irreps=[0, 0, 1, 2, 3, 3])),
    # triplet (na - nb = 2): the non-singlet spin constraints of spin.cc:411-484
    ("triplet_n6", lambda: (Problem(6, 4, 2),) + synthetic_integrals(6, 12, 7)),
]


TIGHT = dict(e_convergence=1e-7, r_convergence=1e-6)
# RDM elements are only determined to the stopping tolerance divided by the
# conditioning of the optimal face: on c2v_n6 the oracle's own 1-RDM moves by
# 1.8e-5 between its LAPACK and Jacobi projections at TIGHT, and by 1.7e-7 at
# TIGHTER, so the 1e-5 RDM comparison runs at TIGHTER.  At TIGHTER the
# oracle's own energies differ by up to 2.2e-8 Eh between those two
# projections (doublet_n5; 8e-9 on c2v_n6, 2.6e-9 on c1_n4): the energy bound
# there is 5e-8.
TIGHTER = dict(e_convergence=1e-9, r_convergence=1e-8)
TIGHTER_ENERGY = 5e-8


def _oracle_solve(prob, h, g, **kw):
    P = OracleProblem(prob.nmo, prob.nalpha, prob.nbeta,
                      prob.irreps if prob.irreps is not None else [0] * prob.nmo, prob.nirrep or 1)
    r = P.solve(P.pack_c(h, g), maxiter=50000, nthreads=4, **kw)
    return P, r


def _gpu_solve(prob, h, g, **kw):
    B = Batch([prob], options(maxiter=50000, **kw))
    B.set_integrals(0, h, g)
    assert B.solve_to_convergence() == 1
    return B


@pytest.mark.parametrize("name,make", SOLVE_CASES)
def test_solve_matches_oracle(name, make):
    prob, h, g = make()
    P, ref = _oracle_solve(prob, h, g, **TIGHT)
    assert ref["converged"]
    B = _gpu_solve(prob, h, g, **TIGHT)
    st = B.status(0)
    assert st["converged"] and st["status_bits"] == 0
    assert abs(st["energy_primal"] - ref["energy_primal"]) < 1e-6
    assert abs(st["energy_primal"] - st["energy_dual"]) < 1e-7
    assert st["eps_p"] < 1e-6 and st["eps_d"] < 1e-6
    B.close()


@pytest.mark.parametrize("name,make", SOLVE_CASES)
def test_rdms_match_oracle(name, make):
    prob, h, g = make()
    P, ref = _oracle_solve(prob, h, g, **TIGHTER)
    assert ref["converged"]
    B = _gpu_solve(prob, h, g, **TIGHTER)
    st = B.status(0)
    assert st["converged"] and st["status_bits"] == 0
    assert abs(st["energy_primal"] - ref["energy_primal"]) < TIGHTER_ENERGY
    x = B.x(0)
    assert np.abs(P.opdm(x) - P.opdm(ref["x"])).max() < 1e-5
    assert abs(P.spin_squared(x) - P.spin_squared(ref["x"])) < 1e-5
    B.close()


@pytest.mark.parametrize("name,make", SOLVE_CASES)
def test_solve_default_thresholds(name, make):
    prob, h, g = make()
    _, ref = _oracle_solve(prob, h, g, **TIGHT)
    B = _gpu_solve(prob, h, g)
    st = B.status(0)
    assert st["converged"] and st["status_bits"] == 0
    assert abs(st["energy_primal"] - st["energy_dual"]) < 1e-6
    assert st["eps_p"] < 1e-5 and st["eps_d"] < 1e-5
    assert abs(st["energy_primal"] - ref["energy_primal"]) < 5e-5
    B.close()


def test_batch_equals_single():
    made = [make() for _, make in SOLVE_CASES]
    B = Batch([m[0] for m in made], options(maxiter=20000))
    for i, (_, h, g) in enumerate(made):
        B.set_integrals(i, h, g)
    assert B.solve_to_convergence() == len(made)
    for i, (prob, h, g) in enumerate(made):
        S = Batch([prob], options(maxiter=20000))
        S.set_integrals(0, h, g)
        S.solve_to_convergence()
        a, b = B.status(i), S.status(0)
        assert a["oiter"] == b["oiter"] and a["iiter"] == b["iiter"]
        assert np.array_equal(B.x(i), S.x(0))
        S.close()
    B.close()


GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_solutions.json")


@pytest.mark.parametrize("name", ["hubbard8", "c1_10"])
def test_configs_against_oracle_golden(name):
    from hilbert_amd import configs
    gold_all = json.load(open(GOLDEN))
    gold = gold_all[name]
    prob, h, g = configs.CONFIGS[name]()[0]
    B = _gpu_solve(prob, h, g, e_convergence=gold_all["e_convergence"],
                   r_convergence=gold_all["r_convergence"])
    st = B.status(0)
    assert abs(st["energy_primal"] - gold["energy_primal"]) < 1e-6
    P = _oracle(prob.nmo, prob.nalpha, prob.nbeta, None, 1)
    x = B.x(0)
    opdm = P.opdm(x)
    assert np.abs(opdm.ravel() - np.asarray(gold["opdm"])).max() < 1e-5
    assert abs(P.spin_squared(x) - gold["spin_squared"]) < 1e-5
    B.close()


BATCH_GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "batch64_d2h_oracle.json")


def test_mixed_batch_matches_oracle_golden():
    """The first 16 instances of the mixed (6e,6o)-(12e,12o) D2h batch, solved
    as one batch, against the oracle's energies of each instance (both at
    1e-9 / 1e-8, where two solvers agree to the TIGHTER_ENERGY bound)."""
    from hilbert_amd import configs
    gold = json.load(open(BATCH_GOLDEN))
    items = configs.batch64_d2h()[:gold["count"]]
    B = Batch([p for p, _, _ in items], options(maxiter=200000, e_convergence=gold["e_convergence"],
                                                 r_convergence=gold["r_convergence"]))
    for i, (_, h, g) in enumerate(items):
        B.set_integrals(i, h, g)
    assert B.solve_to_convergence() == len(items)
    for i, ((prob, _, _), ref) in enumerate(zip(items, gold["instances"])):
        assert list(prob.irreps) == ref["irreps"] and prob.nmo == ref["nmo"]
        st = B.status(i)
        assert st["converged"] and st["status_bits"] == 0
        assert abs(st["energy_primal"] - ref["energy_primal"]) < TIGHTER_ENERGY, \
            (i, st["energy_primal"], ref["energy_primal"])
    B.close()


def _eigh_projection(U, dims):
    """x = U+ per block from LAPACK (numpy eigh) of the symmetrised block."""
    out = np.empty_like(U)
    off = 0
    for d in dims:
        if d == 0:
            continue
        blk = U[off:off + d * d].reshape(d, d)
        w, V = np.linalg.eigh(0.5 * (blk + blk.T))
        out[off:off + d * d] = ((V * np.maximum(w, 0.0)) @ V.T).ravel()
        off += d * d
    return out


PROJ_CASES = {
    # blocks: Hubbard 8 (128, 64, 28, 8), C1 n=10 (200, 100, 45, 10),
    # (16e,16o) C2v (up to about 180), C1 n=11 (242: matrix in global memory)
    "hubbard8": lambda: (Problem(8, 4, 4), *hubbard_integrals(8, 1.0, 4.0)),
    "c1_10": lambda: _synthetic(10, 30, 1, 11),
    "c2v_16": lambda: _synthetic(16, 48, 4, 12),
    "c1_11": lambda: _synthetic(11, 20, 1, 13),
    # C1 n=12 (G2aa 288: Jacobi in global memory); first instance of the
    # large D2h batch, (20e,20o)-(24e,24o) (blocks up to about 180)
    "c1_12": lambda: _synthetic(12, 20, 1, 14),
    "batch16_big_0": lambda: __import__("hilbert_amd.configs", fromlist=["x"]).batch16_big(1)[0],
}


def _synthetic(n, naux, nirrep, seed):
    from hilbert_amd.integrals import random_irreps
    rng = np.random.default_rng(seed)
    irr = random_irreps(n, nirrep, rng) if nirrep > 1 else None
    h, g = synthetic_integrals(n, naux, rng, irreps=irr)
    return (Problem(n, (n + 1) // 2, n // 2, irr.tolist() if irr is not None else None, nirrep), h, g)


@pytest.mark.parametrize("name", list(PROJ_CASES))
def test_projection_matches_eigh(name):
    """Each outer iteration's x = U+ / mu agrees with a LAPACK eigendecomposition
    of the same U (first iteration cold, later ones warm-started)."""
    prob, h, g = PROJ_CASES[name]()
    B = Batch([prob], options(maxiter=100000))
    B.set_integrals(0, h, g)
    dims = B.block_dims(0)
    c = B.c(0)
    for it in range(4):
        x_prev, mu = B.x(0), B.mu(0)
        B.solve(1)
        U = mu * x_prev + B.ATu(0, B.y(0)) - c
        ref = _eigh_projection(U, dims) / mu
        x = B.x(0)
        off = 0
        for d in dims:
            if d == 0:
                continue
            scale = max(np.abs(U[off:off + d * d]).max(), 1e-300) / mu
            err = np.abs(x[off:off + d * d] - ref[off:off + d * d]).max() / scale
            assert err < 1e-11, (name, it, d, err)
            off += d * d
        assert B.status(0)["status_bits"] == 0
    B.close()
