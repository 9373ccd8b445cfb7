"""Pin the oracle's boundary-point solve against what the mathematics fixes.

libsdp (the reference's SDP engine) is not vendored, so its iteration
trajectory is unpinned; only converged results are checked:
* 2-site Hubbard at half filling: E = U/2 - sqrt(U^2/4 + 4 t^2);
* two electrons (one alpha, one beta): DQG is exact for the lowest singlet;
* U = 0: E = sum of the lowest na + nb orbital energies;
* E_DQG <= E_FCI of the lowest singlet for a 4-orbital, 4-electron case;
* the symmetry-blocked solve equals the C1 solve of the same Hamiltonian;
* the projection matches numpy's eigh; Jacobi matches LAPACK;
* the converged point satisfies the stopping criteria and weak duality.
"""
import numpy as np
import pytest

from fci import Fock, lowest_spin_state
from hilbert_amd.integrals import hubbard_integrals, synthetic_integrals
from oracle.pyoracle import OracleProblem, dsyevd_pointer, lib

TOL = dict(e_convergence=1e-8, r_convergence=1e-7, maxiter=50000)


def solve(P, h, g, **kw):
    opts = dict(TOL)
    opts.update(kw)
    return P.solve(P.pack_c(h, g), nthreads=2, **opts)


def test_two_site_hubbard_closed_form():
    t, U = 1.0, 4.0
    h, g = hubbard_integrals(2, t, U)
    r = solve(OracleProblem(2, 1, 1), h, g)
    assert r["converged"]
    assert abs(r["energy_primal"] - (U / 2 - np.sqrt(U * U / 4 + 4 * t * t))) < 1e-6


def test_two_electrons_exact():
    n = 4
    h, g = synthetic_integrals(n, 8, 3)
    F = Fock(n)
    e_fci, _ = lowest_spin_state(F, F.hamiltonian(h, g), 1, 1)
    r = solve(OracleProblem(n, 1, 1), h, g)
    assert r["converged"]
    assert abs(r["energy_primal"] - e_fci) < 1e-5


def test_noninteracting_limit():
    n, na = 4, 2
    h, g = hubbard_integrals(n, 1.0, 0.0)
    r = solve(OracleProblem(n, na, na), h, g)
    e = 2 * np.sort(np.linalg.eigvalsh(h))[:na].sum()
    assert r["converged"]
    assert abs(r["energy_primal"] - e) < 1e-5


def test_lower_bound_on_fci():
    n, na = 4, 2
    h, g = synthetic_integrals(n, 8, 9)
    F = Fock(n)
    e_fci, _ = lowest_spin_state(F, F.hamiltonian(h, g), na, na)
    r = solve(OracleProblem(n, na, na), h, g)
    assert r["converged"]
    assert r["energy_primal"] <= e_fci + 1e-6
    assert e_fci - r["energy_primal"] < 0.1  # DQG is tight for small molecules


def test_symmetry_blocked_equals_c1():
    n, na = 4, 2
    irr = [0, 0, 1, 1]
    h, g = synthetic_integrals(n, 8, 4, irreps=irr)
    r_sym = solve(OracleProblem(n, na, na, irr, 2), h, g)
    r_c1 = solve(OracleProblem(n, na, na), h, g)
    assert r_sym["converged"] and r_c1["converged"]
    assert abs(r_sym["energy_primal"] - r_c1["energy_primal"]) < 1e-5
    assert OracleProblem(n, na, na, irr, 2).n_primal < OracleProblem(n, na, na).n_primal


def test_stopping_criteria_and_weak_duality():
    h, g = synthetic_integrals(4, 8, 1)
    P = OracleProblem(4, 2, 2)
    r = P.solve(P.pack_c(h, g), nthreads=2)  # default thresholds 1e-6 / 1e-5
    assert r["converged"]
    assert abs(r["energy_primal"] - r["energy_dual"]) < 1e-6
    assert r["eps_p"] < 1e-5 and r["eps_d"] < 1e-5
    x = r["x"]
    b = P.b()
    assert np.linalg.norm(P.Au(x) - b) / np.linalg.norm(b) == pytest.approx(r["eps_p"], rel=1e-6)
    # every primal block is PSD, every dual slack block too
    off = 0
    for d in P.dims:
        d = int(d)
        blk = x[off:off + d * d].reshape(d, d)
        zb = r["z"][off:off + d * d].reshape(d, d)
        assert np.linalg.eigvalsh(blk).min() > -1e-10
        assert np.linalg.eigvalsh(zb).min() > -1e-10
        off += d * d


@pytest.mark.parametrize("use_lapack", [True, False])
def test_projection_matches_eigh(use_lapack):
    rng = np.random.default_rng(0)
    for d in (1, 3, 17, 40):
        U = rng.standard_normal((d, d))
        mu = 0.7
        lw = 8 * d * d + 16 * d + 128
        work = np.zeros(lw)
        iwork = np.zeros(5 * d + 32, dtype=np.int32)
        x = np.zeros((d, d))
        z = np.zeros((d, d))
        eig = dsyevd_pointer() if use_lapack else None
        rc = lib().orc_project_block(d, np.ascontiguousarray(U), mu, x, z, eig, work, lw, iwork,
                                     len(iwork))
        assert rc == 0
        S = 0.5 * (U + U.T)
        w, v = np.linalg.eigh(S)
        xp = (v * np.maximum(w, 0)) @ v.T / mu
        zp = -(v * np.minimum(w, 0)) @ v.T
        assert np.abs(x - xp).max() < 1e-12 * max(1, np.abs(S).max())
        assert np.abs(z - zp).max() < 1e-12 * max(1, np.abs(S).max())
