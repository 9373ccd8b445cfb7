"""Solver contract of libsdp's BPSDPSolver as hilbert uses it
(v2rdm_solver.cc:1305-1361, checkpoint.cc:58-120): state persists across
solve() calls, MAXITER caps the outer iterations, the monitor sees every
outer iteration, and a checkpoint (x, y, z, mu) restores a solve."""
import numpy as np
import pytest

from hilbert_amd.batch import Batch, Problem, options
from hilbert_amd.integrals import hubbard_integrals, synthetic_integrals

pytestmark = pytest.mark.gpu


def _cases():
    return [(Problem(2, 1, 1),) + hubbard_integrals(2, 1.0, 4.0),
            (Problem(4, 2, 2),) + synthetic_integrals(4, 8, 1),
            (Problem(5, 3, 2),) + synthetic_integrals(5, 10, 2)]


def _batch(items, **kw):
    B = Batch([p for p, _, _ in items], options(**kw))
    for i, (_, h, g) in enumerate(items):
        B.set_integrals(i, h, g)
    return B


def test_chunked_solves_equal_one_call():
    """solve(a) then solve(n - a) gives the bits of solve(n): y, z, mu and the
    eigenvalue warm start all persist between calls."""
    items = _cases()
    one = _batch(items, maxiter=100000)
    one.solve(300)
    two = _batch(items, maxiter=100000)
    two.solve(117)
    two.solve(183)
    for i in range(len(items)):
        a, b = one.status(i), two.status(i)
        assert a["oiter"] == b["oiter"] and a["iiter"] == b["iiter"]
        assert np.array_equal(one.x(i), two.x(i))
        assert np.array_equal(one.y(i), two.y(i))
        assert one.mu(i) == two.mu(i)
    one.close()
    two.close()


def test_maxiter_caps_outer_iterations():
    items = _cases()
    B = _batch(items, maxiter=25)
    B.solve(1000)
    for i in range(len(items)):
        st = B.status(i)
        assert st["oiter"] == 25 and not st["converged"] and not st["active"]
    B.solve(1000)  # nothing left to run
    assert all(B.status(i)["oiter"] == 25 for i in range(len(items)))
    B.close()


def test_monitor_sees_every_outer_iteration():
    items = _cases()
    B = _batch(items, maxiter=100000)
    seen = {i: [] for i in range(len(items))}

    def mon(inst, oiter, iiter, ep, ed, mu, errp, errd):
        seen[inst].append((oiter, ep, ed))

    B.solve(40, monitor=mon)
    for i in range(len(items)):
        assert [o for o, _, _ in seen[i]] == list(range(1, 41))
        st = B.status(i)
        assert seen[i][-1][1] == st["energy_primal"] and seen[i][-1][2] == st["energy_dual"]
    B.close()


def test_checkpoint_restores_a_solve():
    """Stop after 200 iterations, copy x, y, z and mu into a new handle and
    finish there: the energy agrees with an uninterrupted solve."""
    items = _cases()
    ref = _batch(items, maxiter=100000, e_convergence=1e-7, r_convergence=1e-6)
    assert ref.solve_to_convergence() == len(items)
    A = _batch(items, maxiter=100000, e_convergence=1e-7, r_convergence=1e-6)
    A.solve(200)
    B = _batch(items, maxiter=100000, e_convergence=1e-7, r_convergence=1e-6)
    for i in range(len(items)):
        B.set_guess(i, A.x(i))
        B.set_y(i, A.y(i))
        B.set_z(i, A.z(i))
        B.set_mu(i, A.mu(i))
    assert B.solve_to_convergence() == len(items)
    for i in range(len(items)):
        assert abs(B.status(i)["energy_primal"] - ref.status(i)["energy_primal"]) < 1e-6
    for h in (ref, A, B):
        h.close()


def test_restored_state_continues_the_iteration():
    """x, y, z and mu copied into a fresh handle continue the same iteration:
    the next outer iteration from the restored state equals the next outer
    iteration of the original handle (A^T y is rebuilt with y, so the CG's
    first residual is that of the copied state; only the projection's
    eigenvalue warm start differs, which moves the last bits).  CG_CONVERGENCE
    1e-2 gives both handles the same CG tolerance (a fresh handle starts from
    errors of 1)."""
    items = _cases()
    A = _batch(items, maxiter=100000, cg_convergence=1e-2)
    A.solve(150)
    B = _batch(items, maxiter=100000, cg_convergence=1e-2)
    for i in range(len(items)):
        B.set_guess(i, A.x(i))
        B.set_y(i, A.y(i))
        B.set_z(i, A.z(i))
        B.set_mu(i, A.mu(i))
    A.solve(1)
    B.solve(1)
    for i in range(len(items)):
        ya, yb = A.y(i), B.y(i)
        assert np.abs(ya - yb).max() <= 1e-9 * max(1.0, np.abs(ya).max()), i
        assert np.abs(A.x(i) - B.x(i)).max() <= 1e-9, i
    A.close()
    B.close()
