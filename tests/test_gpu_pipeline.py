"""GPU: the pipelined schedule of the batched solve (v2h_tuning.pipeline).

With pipelining the running instances form two groups that take turns on the
main stream (pre, CG, post) while the other group's projection runs on the
side streams.  Every instance goes through the same kernels on the same data
in the same order as in lock-step, so the two schedules must agree bit for
bit:

* a mixed batch solved to convergence: its instances stop at different
  iterations, so the groups are split again when they become unbalanced and
  the batch drops to lock-step when fewer instances than the threshold run;
* the same with the tridiagonal projection only (no refinement);
* a batch solved in short calls, switching the schedule between calls;
* with the kernel timers on (their events on streams that the main stream
  has not joined yet are read later).
"""
import numpy as np
import pytest

from hilbert_amd import configs
from hilbert_amd.batch import Batch, options

pytestmark = pytest.mark.gpu

OPTS = dict(e_convergence=1e-7, r_convergence=1e-6, maxiter=100000)


def _items():
    return configs.c2v(16, 4) + configs.batch64_d2h(6)


def _batch(items, pipeline, refine=True, profile=False, opts=OPTS):
    B = Batch([p for p, _, _ in items], options(**opts))
    B.set_tuning(dict(B.tuning(), pipeline=pipeline))
    assert B.tuning()["pipeline"] == pipeline
    B.set_projection(Batch.PROJECT_REFINED if refine else Batch.PROJECT_TRIDIAGONAL)
    B.set_profiling(profile)
    for i, (_, h, g) in enumerate(items):
        B.set_integrals(i, h, g)
    return B


def _result(B):
    return [(B.x(i), B.y(i), B.z(i), B.status(i)) for i in range(B.count)]


def _same(a, b):
    for i, (ra, rb) in enumerate(zip(a, b)):
        assert ra[3]["oiter"] == rb[3]["oiter"] and ra[3]["iiter"] == rb[3]["iiter"], (i, ra[3], rb[3])
        assert ra[3]["converged"] == rb[3]["converged"], i
        for k in range(3):
            assert np.array_equal(ra[k], rb[k]), (i, k)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("refine", [True, False])
def test_pipelined_solve_equals_lock_step(refine):
    items = _items()
    res = []
    for pipeline in (0, 4):
        B = _batch(items, pipeline, refine)
        B.solve_to_convergence()
        res.append(_result(B))
        B.close()
    oiters = sorted({r[3]["oiter"] for r in res[0]})
    print(f"stop iterations {oiters}", flush=True)
    assert len(oiters) > 1
    assert all(r[3]["converged"] for r in res[0])
    _same(res[1], res[0])


@pytest.mark.timeout(600)
def test_switching_schedules_between_short_calls():
    items = configs.c2v(16, 4)
    ref = _batch(items, 0)
    ref.solve_to_convergence()
    want = _result(ref)
    ref.close()
    B = _batch(items, 2)
    B.solve(150)
    for k in range(20):
        B.set_tuning(dict(B.tuning(), pipeline=2 if k % 2 else 0))
        B.solve(7 + k % 3)
    B.set_tuning(dict(B.tuning(), pipeline=2))
    B.solve_to_convergence()
    _same(_result(B), want)
    B.close()


def test_kernel_timers_with_pipelining():
    items = _items()
    never = dict(e_convergence=1e-14, r_convergence=1e-13, maxiter=1000000)
    out = []
    for pipeline in (0, 4):
        B = _batch(items, pipeline, profile=True, opts=never)
        B.solve(120)
        out.append((_result(B), B.kernel_times()))
        B.close()
    (ra, ta), (rb, tb) = out
    _same(rb, ra)
    # the same work in twice the launches of the main-stream kernels (one per group)
    for kind in ("pre", "post", "project_refine", "project_seed"):
        assert tb[kind][1] >= ta[kind][1] > 0, (kind, ta[kind], tb[kind])
        assert tb[kind][0] > 0.0, kind
