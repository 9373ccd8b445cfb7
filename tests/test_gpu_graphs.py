"""GPU: hipGraph replay of the outer iteration's fixed launch sequences.

Mode GRAPHS_CG captures the first CG chunk of every outer iteration, mode
GRAPHS_ALL also k_post + the whole projection (the eigenvector formation then
runs as a branch of the next CG graph); each graph is captured once per
launch plan and replayed.  Without graphs every kernel is launched one by
one.  The kernels are the same and deterministic, so all modes must agree
bit for bit:

* a batch of two (16e,16o) instances and two small ones, 400 outer iterations
  (hundreds of replays, interleaved with eager launches and host polls), with
  the refinement on and off;
* a mixed batch solved to convergence, whose instances stop at different
  iterations, so the plan shrinks and the graphs are captured again;
* with the kernel timers on, the graphs' event nodes give the same launch
  counts as the eager launches.
"""
import numpy as np
import pytest

from hilbert_amd import configs
from hilbert_amd.batch import Batch, options

pytestmark = pytest.mark.gpu

NEVER = dict(e_convergence=1e-14, r_convergence=1e-13, maxiter=1000000)


def _items():
    return configs.c2v(16, 2) + configs.batch64_d2h(2)


def _run(items, graphs, refine, iters, opts, profile=False):
    B = Batch([p for p, _, _ in items], options(**opts))
    B.set_graphs(graphs)
    assert B.tuning()["graphs"] == graphs
    B.set_projection(Batch.PROJECT_REFINED if refine else Batch.PROJECT_TRIDIAGONAL)
    B.set_profiling(profile)
    for i, (_, h, g) in enumerate(items):
        B.set_integrals(i, h, g)
    if iters is None:
        B.solve_to_convergence()
    else:
        B.solve(iters)
    out = [(B.x(i), B.y(i), B.z(i), B.status(i)) for i in range(len(items))]
    times = B.kernel_times() if profile else None
    B.close()
    return out, times


@pytest.mark.parametrize("mode", [Batch.GRAPHS_CG, Batch.GRAPHS_ALL])
@pytest.mark.parametrize("refine", [True, False])
def test_graphs_equal_eager_launches(refine, mode):
    items = _items()
    a, _ = _run(items, mode, refine, 400, NEVER)
    b, _ = _run(items, Batch.GRAPHS_OFF, refine, 400, NEVER)
    for i, (ra, rb) in enumerate(zip(a, b)):
        assert ra[3]["oiter"] == rb[3]["oiter"] and ra[3]["iiter"] == rb[3]["iiter"], i
        for k in range(3):
            assert np.array_equal(ra[k], rb[k]), (i, k)


def test_graphs_follow_a_shrinking_plan():
    items = configs.batch64_d2h(5) + configs.c2v(16, 1)
    opts = dict(e_convergence=1e-6, r_convergence=1e-5, maxiter=100000)
    a, ta = _run(items, Batch.GRAPHS_ALL, True, None, opts, profile=True)
    b, tb = _run(items, Batch.GRAPHS_OFF, True, None, opts, profile=True)
    oiters = sorted({r[3]["oiter"] for r in a})
    print(f"stop iterations {oiters}", flush=True)
    assert len(oiters) > 1  # instances stopped at different iterations
    for i, (ra, rb) in enumerate(zip(a, b)):
        assert ra[3]["converged"] and ra[3]["oiter"] == rb[3]["oiter"], i
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1]), i
    for kind in ("pre", "cg_at", "cg_a", "post", "project_refine", "project_seed"):
        assert ta[kind][1] == tb[kind][1] > 0, (kind, ta[kind], tb[kind])
        assert ta[kind][0] > 0.0
