"""GPU: v2h_batch_Au_noalloc, A u staged through the library's static device
scratch instead of device allocations, through the C ABI.

It must give v2h_batch_Au's bits on two calls in a row (the second with a
smaller instance, so that the scratch still holds the first call's longer
vectors), again on the first instance afterwards, and reject an instance
whose vectors do not fit the scratch.
"""
import numpy as np
import pytest

from hilbert_amd import _native as N
from hilbert_amd import configs
from hilbert_amd.batch import Batch, options

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 1 << 19


def _au_noalloc(B, i, u):
    out = np.full(B.n_dual[i], np.nan)
    N.check(N.lib().v2h_batch_Au_noalloc(B._h, i, np.ascontiguousarray(u, dtype=np.float64), out))
    return out


def test_au_noalloc_matches_au():
    # (16e,16o) C2v, then a smaller (6e,6o)-(12e,12o) D2h instance
    items = [configs.c2v(16, 1)[0], configs.batch64_d2h(1)[0]]
    B = Batch([p for p, _, _ in items], options())
    for i, (_, h, g) in enumerate(items):
        B.set_integrals(i, h, g)
    assert B.n_primal[1] + B.n_dual[1] < B.n_primal[0] + B.n_dual[0] <= SCRATCH_DOUBLES
    rng = np.random.default_rng(7)
    us = [rng.standard_normal(B.n_primal[i]) for i in range(2)]
    print(f"noalloc: instance sizes {B.n_primal} primal, {B.n_dual} dual", flush=True)
    got = [_au_noalloc(B, i, us[i]) for i in range(2)]  # two calls in a row, the second smaller
    for i in range(2):
        ref = B.Au(i, us[i])
        assert np.array_equal(got[i], ref), i
    assert np.array_equal(_au_noalloc(B, 0, us[0]), got[0])
    B.close()


def test_au_noalloc_rejects_an_instance_beyond_the_scratch():
    prob, h, g = configs.batch16_big(2)[1]  # (24e,24o) D2h
    B = Batch([prob], options())
    assert B.n_primal[0] + B.n_dual[0] > SCRATCH_DOUBLES
    u = np.zeros(B.n_primal[0])
    out = np.zeros(B.n_dual[0])
    rc = N.lib().v2h_batch_Au_noalloc(B._h, 0, u, out)
    assert rc == 1 and b"scratch" in N.lib().v2h_last_error()
    B.close()
