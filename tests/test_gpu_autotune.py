"""GPU: the launch-configuration autotuner (hilbert_amd/autotune.py).

Its knobs (graph replay; stream of each PSD size class, of the refinement and
of its fallback; the CG chunk margin; the XCD workgroup remap) change no
result bit, so a solve tuned on the fly returns the bits of an untuned one.
Checked: the
tuner times every candidate, keeps a valid configuration, writes it to the
cache under the batch signature, and a second batch with the same signature
picks it up; the tuned solve equals the untuned one bit for bit.
"""
import json

import numpy as np
import pytest

from hilbert_amd import autotune, configs
from hilbert_amd.batch import Batch, options

pytestmark = pytest.mark.gpu

OPTS = dict(e_convergence=1e-7, r_convergence=1e-6, maxiter=100000)


def _batch(items):
    B = Batch([p for p, _, _ in items], options(**OPTS))
    for i, (_, h, g) in enumerate(items):
        B.set_integrals(i, h, g)
    return B


@pytest.mark.timeout(600)
def test_autotuned_solve_equals_untuned_bits(tmp_path):
    items = configs.c2v(16, 4)
    cache = str(tmp_path / "tuning.json")
    A = _batch(items)
    A.solve(300)
    out = autotune.tune(A, window=10, rounds=2, cache=cache, log=print)
    cfg = out["tuning"]
    assert set(out["ms_per_iteration"]) == {"graphs", "xcd_remap", "cg_margin", "lanes"}
    assert all(v is not None for knob in out["ms_per_iteration"].values() for v in knob.values())
    assert A.tuning() == cfg
    A.solve_to_convergence()
    ref = _batch(items)
    ref.solve_to_convergence()
    for i in range(len(items)):
        sa, sr = A.status(i), ref.status(i)
        assert sa["converged"] and sa["oiter"] == sr["oiter"] and sa["iiter"] == sr["iiter"], i
        assert np.array_equal(A.x(i), ref.x(i)) and np.array_equal(A.y(i), ref.y(i)), i
    data = json.load(open(cache))
    assert data[autotune.signature(A)]["tuning"] == cfg
    C = _batch(items)
    assert autotune.apply_cached(C, cache) == cfg and C.tuning() == cfg
    for B in (A, ref, C):
        B.close()


def test_tuning_validation():
    prob, h, g = configs.batch64_d2h(1)[0]
    B = Batch([prob], options())
    cfg = B.tuning()
    with pytest.raises(Exception):
        B.set_tuning(dict(cfg, refine_lane=0))
    with pytest.raises(Exception):
        B.set_tuning(dict(cfg, class_lane=[0, 0, 4, 0, 0, 0, 0]))
    assert B.tuning() == cfg
    B.close()
