"""GPU: the whole batched workload of 64 independent SDPs, (6e,6o) to
(12e,12o) with D2h labels at random, solved as one batch and checked per
instance against the oracle's energies (tests/golden/batch64_d2h_all_oracle.json,
oracle BPSDP at 1e-9 / 1e-8).

At 1e-8 / 1e-7 every instance must converge and land within 1e-6 Eh of the
oracle's optimum; the instances stop one by one while the others keep
iterating, so this also exercises the shrinking launch plans over uneven
block lists.
"""
import json
import os

import numpy as np
import pytest

from hilbert_amd import configs
from hilbert_amd.batch import Batch, options

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_all_64_d2h_instances_reach_the_oracle_energy():
    items = configs.batch64_d2h(64)
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "batch64_d2h_all_oracle.json")))
    ref = {o["index"]: o for o in fix["instances"]}
    assert len(ref) == 64
    B = Batch([p for p, _, _ in items], options(e_convergence=1e-8, r_convergence=1e-7, maxiter=100000))
    for i, (p, h, g) in enumerate(items):
        assert (p.nmo, p.nalpha, p.nbeta) == (ref[i]["nmo"], ref[i]["nalpha"], ref[i]["nbeta"])
        B.set_integrals(i, h, g)
    B.solve_to_convergence()
    st = [B.status(i) for i in range(64)]
    B.close()
    dE = np.array([abs(s["energy_primal"] - ref[i]["energy_primal"]) for i, s in enumerate(st)])
    print(f"64 x (6e-12e) D2h: {sum(s['converged'] for s in st)} converged, max |dE| {dE.max():.2e} "
          f"(instance {int(dE.argmax())}), outer iterations {min(s['oiter'] for s in st)}.."
          f"{max(s['oiter'] for s in st)}", flush=True)
    assert all(s["converged"] and s["status_bits"] == 0 for s in st)
    assert dE.max() <= 1e-6, list(np.nonzero(dE > 1e-6)[0])
