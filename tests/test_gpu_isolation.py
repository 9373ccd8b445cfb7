"""GPU: per-instance failure isolation in a batch.

The reference's loop has no exit besides is_converged()
(v2rdm_solver.cc:1330-1361); a batch must not let one failing SDP stall or
change the others.  One instance gets NaN integrals (caught when its cost
vector is packed, before any kernel iterates it) or a NaN starting point
(caught by the device-side book-keeping after its first outer iteration).
In both cases that instance stops with the non-finite status bit and every
other instance returns the same bits as when it is solved alone.
"""
import numpy as np
import pytest

from hilbert_amd.batch import Batch, Problem, options
from hilbert_amd.integrals import synthetic_integrals

pytestmark = pytest.mark.gpu

NONFINITE = 2  # kStatusNonFinite
OPTS = dict(maxiter=20000)


def _items():
    out = []
    for k, irr in enumerate(([0, 0, 1, 2, 3, 3], [0, 1, 1, 2, 2, 3], [0, 0, 1, 2, 3])):
        n = len(irr)
        rng = np.random.default_rng(500 + k)
        h, g = synthetic_integrals(n, 2 * n, rng, irreps=irr)
        out.append((Problem(n, (n + 1) // 2, n // 2, irr, 4), h, g))
    return out


def _solo(p, h, g):
    B = Batch([p], options(**OPTS))
    B.set_integrals(0, h, g)
    B.solve_to_convergence()
    out = B.status(0), B.x(0), B.y(0)
    B.close()
    return out


@pytest.mark.parametrize("where", ["integrals", "guess"])
def test_nan_instance_is_isolated(where):
    items = _items()
    bad = 1  # the poisoned instance sits between healthy ones
    p0, h0, g0 = items[0]
    batch = [items[0], (p0, h0.copy(), g0.copy()), items[1], items[2]]
    if where == "integrals":
        batch[bad][1][0, 0] = np.nan
    print(f"isolation case: NaN in the {where} of instance {bad}", flush=True)
    B = Batch([p for p, _, _ in batch], options(**OPTS))
    for i, (_, h, g) in enumerate(batch):
        B.set_integrals(i, h, g)
    if where == "guess":
        x = B.x(bad)
        x[0] = np.nan
        B.set_guess(bad, x)
    B.solve_to_convergence()
    st = [B.status(i) for i in range(B.count)]
    xs = [B.x(i) for i in range(B.count)]
    ys = [B.y(i) for i in range(B.count)]
    B.close()

    s = st[bad]
    assert s["status_bits"] & NONFINITE, s
    assert not s["converged"] and not s["active"]
    assert s["oiter"] == (0 if where == "integrals" else 1), s
    for i in (0, 2, 3):
        ref_st, ref_x, ref_y = _solo(*batch[i])
        assert st[i]["converged"] and st[i]["status_bits"] == 0
        assert st[i]["oiter"] == ref_st["oiter"] and st[i]["iiter"] == ref_st["iiter"]
        assert st[i]["energy_primal"] == ref_st["energy_primal"]
        assert np.array_equal(xs[i], ref_x) and np.array_equal(ys[i], ref_y)
