"""GPU: the differentiable energy at (16e,16o) C2v (naux 48) against central
finite differences, step 1e-5, on 50 random integral elements (25 of h, 25
of (pq|rs)), and the 1-RDM it implies against the oracle's.

dE/dtheta comes from v2rdm_energies (the torch.autograd.Function over the
adjoint of the cost packing, envelope theorem).  Each difference quotient
needs two energies accurate to about 1e-11 Eh: the 100 perturbed SDPs are
solved as one batch, warm-started from the converged unperturbed state
(x, y, z, mu), to E_CONVERGENCE 1e-12 / R_CONVERGENCE 1e-11, then 2000 more
outer iterations towards 1e-13 / 1e-12, with CG_CONVERGENCE 1e-14 (at
hilbert's default CG_CONVERGENCE 1e-9 the CG floor stalls BPSDP near 1e-10
relative residuals).

dE/dh_ij = D1a_ij + D1b_ij (c[D1a] = c[D1b] = h, v2rdm_solver.cc:164-170), so
the h-gradient is the spin-summed 1-RDM; it is compared with the oracle's
1-RDM of the same instance (tests/golden/c2v16_oracle.json, 1e-9 / 1e-8).
"""
import json
import os

import numpy as np
import pytest
import torch

from hilbert_amd import configs
from hilbert_amd.adjoint import v2rdm_energies_problems
from hilbert_amd.batch import Batch, options

pytestmark = pytest.mark.gpu

STEP = 1e-5
TIGHT = dict(E_CONVERGENCE=1e-11, R_CONVERGENCE=1e-10, CG_CONVERGENCE=1e-14, MAXITER=200000)
FD_OPTS = dict(e_convergence=1e-12, r_convergence=1e-11, cg_convergence=1e-14, maxiter=200000)


@pytest.mark.timeout(900)
def test_adjoint_gradient_matches_central_differences_at_16e16o():
    prob, h, g = configs.c2v(16, 1)[0]
    n = prob.nmo
    ht = torch.tensor(h, dtype=torch.float64, requires_grad=True)
    gt = torch.tensor(g, dtype=torch.float64, requires_grad=True)
    info = {}
    E = v2rdm_energies_problems([prob], [ht], [gt], TIGHT, info)
    E.sum().backward()
    dh, dg = ht.grad.numpy(), gt.grad.numpy()
    assert info["status"][0]["converged"]
    e0 = float(E[0])

    # the unperturbed state, for the warm starts (same options and data: the
    # same solve as inside v2rdm_energies, bit for bit)
    B = Batch([prob], options(e_convergence=1e-11, r_convergence=1e-10, cg_convergence=1e-14, maxiter=200000))
    B.set_integrals(0, h, g)
    B.solve_to_convergence()
    assert B.status(0)["energy_primal"] == e0
    x, y, z, mu = B.x(0), B.y(0), B.z(0), B.mu(0)
    B.close()

    rng = np.random.default_rng(1605)
    hi = np.argwhere(np.abs(dh) > 1e-6)  # symmetry-allowed elements (the others do not enter c)
    gi = np.argwhere(np.abs(dg) > 1e-6)
    els = [("h", tuple(hi[k])) for k in rng.choice(len(hi), 25, replace=False)] + \
          [("g", tuple(gi[k])) for k in rng.choice(len(gi), 25, replace=False)]
    P = Batch([prob] * (2 * len(els)), options(**FD_OPTS))
    for k, (kind, idx) in enumerate(els):
        for s, sgn in enumerate((1.0, -1.0)):
            hh, gg = h.copy(), g.copy()
            (hh if kind == "h" else gg)[idx] += sgn * STEP
            i = 2 * k + s
            P.set_integrals(i, hh, gg)
            P.set_guess(i, x)
            P.set_y(i, y)
            P.set_z(i, z)
            P.set_mu(i, mu)
    P.solve_to_convergence()
    conv = [P.status(i)["converged"] for i in range(P.count)]
    # then 2000 more outer iterations towards 1e-13 / 1e-12, from where the
    # first stage stopped (bounded: near the FP64 floor it may not get there)
    P.set_options(options(**dict(FD_OPTS, e_convergence=1e-13, r_convergence=1e-12)))
    P.solve(2000)
    st = [P.status(i) for i in range(P.count)]
    P.close()
    Ep = np.array([s["energy_primal"] for s in st])
    fd = (Ep[0::2] - Ep[1::2]) / (2 * STEP)
    grad = np.array([dh[idx] if kind == "h" else dg[idx] for kind, idx in els])
    err = np.abs(fd - grad)
    print(f"(16e,16o) FD check: {sum(conv)}/{len(st)} perturbed solves converged to 1e-12 / 1e-11, "
          f"{sum(s['converged'] for s in st)} to 1e-13 / 1e-12, "
          f"max outer iterations {max(s['oiter'] for s in st)}; |FD - adjoint| max {err.max():.2e} "
          f"mean {err.mean():.2e}; |grad| range {np.abs(grad).min():.1e}..{np.abs(grad).max():.1e}", flush=True)
    assert all(conv)
    assert err.max() < 1e-6, list(zip(els, err))

    # dE/dh = the spin-summed 1-RDM, against the oracle's
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "c2v16_oracle.json")))
    ref = {o["index"]: o for o in fix["instances"]}[0]
    d1 = np.array(ref["opdm"]).reshape(n, n)
    assert abs(e0 - ref["energy_primal"]) < 1e-7
    assert np.abs(dh - d1).max() < 1e-5, np.abs(dh - d1).max()
