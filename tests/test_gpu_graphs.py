"""GPU: hipGraph replay of the outer iteration's first CG chunk.

Mode GRAPHS_CG captures the first CG chunk of every outer iteration (its
k_cg_at / k_cg_a / k_cg_reduce launches) once per launch plan and step count
and replays it (with the kernel timers off; with them on the launches go one
by one); GRAPHS_ALL also captures k_post and the whole projection, with the
eigenvector formation as a branch of the next CG graph.  Without graphs
every kernel is launched one by one.  The kernels are the same and
deterministic, so all three must agree bit for bit:

* a batch of two (16e,16o) instances and two small ones, 400 outer iterations
  (hundreds of replays, interleaved with eager launches and host polls), with
  the refinement on and off;
* a mixed batch solved to convergence, whose instances stop at different
  iterations, so the plan shrinks and the graphs are captured again;
* a batch solved in calls of 11 outer iterations, switching the mode at
  every call (GRAPHS_ALL three times over: its eigenvector formation runs
  after the last iteration's bookkeeping at the end of a call, and must
  form the same eigenvectors as the eager launch order).
"""
import numpy as np
import pytest

from hilbert_amd import configs
from hilbert_amd._native import V2hError
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


@pytest.mark.parametrize("mode", [Batch.GRAPHS_CG, Batch.GRAPHS_ALL])
def test_graphs_follow_a_shrinking_plan(mode):
    items = configs.batch64_d2h(5) + configs.c2v(16, 1)
    opts = dict(e_convergence=1e-6, r_convergence=1e-5, maxiter=100000)
    a, _ = _run(items, mode, True, None, opts)
    b, _ = _run(items, Batch.GRAPHS_OFF, True, None, opts)
    oiters = sorted({r[3]["oiter"] for r in a})
    print(f"stop iterations {oiters}", flush=True)
    assert len(oiters) > 1  # instances stopped at different iterations
    for i, (ra, rb) in enumerate(zip(a, b)):
        assert ra[3]["converged"] and ra[3]["oiter"] == rb[3]["oiter"], i
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1]), i


@pytest.mark.parametrize("mode,repeats", [(Batch.GRAPHS_CG, 1), (Batch.GRAPHS_ALL, 3)])
def test_switching_graphs_between_short_calls(mode, repeats):
    items = configs.c2v(16, 4)
    opts = dict(e_convergence=1e-7, r_convergence=1e-6, maxiter=100000)
    ref, _ = _run(items, Batch.GRAPHS_OFF, True, None, opts)
    for rep in range(repeats):
        B = Batch([p for p, _, _ in items], options(**opts))
        for i, (_, h, g) in enumerate(items):
            B.set_integrals(i, h, g)
        B.solve(300)
        for k in range(20):
            B.set_graphs(mode if k % 2 == 0 else Batch.GRAPHS_OFF)
            B.solve(11)
        B.set_graphs(mode)
        B.solve_to_convergence()
        for i, r in enumerate(ref):
            st = B.status(i)
            assert st["oiter"] == r[3]["oiter"] and st["iiter"] == r[3]["iiter"], (rep, i)
            assert np.array_equal(B.x(i), r[0]), (rep, i)
        B.close()


def test_unknown_graph_mode_is_rejected():
    prob, h, g = configs.batch64_d2h(1)[0]
    B = Batch([prob], options())
    with pytest.raises(V2hError):
        B.set_graphs(3)
    assert B.tuning()["graphs"] == Batch.GRAPHS_OFF
    B.close()


def test_kernel_timers_with_graph_mode():
    """With the timers on, mode 1 launches one by one: the same launch counts
    and bits as mode 0."""
    items = _items()
    a, ta = _run(items, Batch.GRAPHS_CG, True, 200, NEVER, profile=True)
    b, tb = _run(items, Batch.GRAPHS_OFF, True, 200, NEVER, profile=True)
    for kind in ("pre", "cg_at", "cg_a", "post", "project_refine"):
        assert ta[kind][1] == tb[kind][1] > 0, (kind, ta[kind], tb[kind])
    for ra, rb in zip(a, b):
        assert np.array_equal(ra[0], rb[0])
