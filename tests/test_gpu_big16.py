"""GPU: the large batch with gradients -- 16 independent (20e,20o)-(24e,24o)
DQG SDPs with D2h labels at random (naux = 3 n) solved as one batch through
v2rdm_energies, with the batched adjoint gradients of all 16.

Checked per instance: convergence, the energy against the oracle's
(tests/golden/big16_oracle.json, oracle BPSDP at 1e-9 / 1e-8), the 1-RDM
implied by dE/dh against the oracle's 1-RDM, and Euler's identity for the
gradient: c is linear in (h, g) and the constraints do not depend on them, so
E(t h, t g) = t E(h, g) and <h, dE/dh> + <g, dE/dg> = E.
"""
import json
import os

import numpy as np
import pytest
import torch

from hilbert_amd import configs
from hilbert_amd.adjoint import v2rdm_energies_problems

pytestmark = pytest.mark.gpu

OPTS = dict(E_CONVERGENCE=1e-8, R_CONVERGENCE=1e-7, MAXITER=40000)


@pytest.mark.timeout(900)
def test_big16_batch_energies_and_gradients():
    items = configs.batch16_big()
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "big16_oracle.json")))
    ref = {o["index"]: o for o in fix["instances"]}
    hs = [torch.tensor(h, dtype=torch.float64, requires_grad=True) for _, h, _ in items]
    gs = [torch.tensor(g, dtype=torch.float64, requires_grad=True) for _, _, g in items]
    info = {}
    E = v2rdm_energies_problems([p for p, _, _ in items], hs, gs, OPTS, info)
    E.sum().backward()
    worst_e, worst_d = 0.0, 0.0
    for i, (p, h, g) in enumerate(items):
        st = info["status"][i]
        assert st["converged"] and st["status_bits"] == 0, (i, st)
        e = float(E[i])
        dh, dg = hs[i].grad.numpy(), gs[i].grad.numpy()
        euler = float(np.sum(h * dh) + np.sum(g * dg))
        assert abs(euler - e) <= 1e-9 * abs(e), (i, euler, e)
        if i in ref:
            worst_e = max(worst_e, abs(e - ref[i]["energy_primal"]))
            assert abs(e - ref[i]["energy_primal"]) <= 1e-6, (i, e, ref[i]["energy_primal"])
            d1 = np.array(ref[i]["opdm"]).reshape(p.nmo, p.nmo)
            worst_d = max(worst_d, float(np.abs(dh - d1).max()))
    print(f"16 x (20e-24e) D2h batch: {len(ref)} instances with oracle fixtures, max |dE| {worst_e:.2e}, "
          f"max |dE/dh - oracle 1-RDM| {worst_d:.2e}; outer iterations "
          f"{[s['oiter'] for s in info['status']]}", flush=True)
    assert worst_d < 1e-4
