"""Differentiable v2RDM energies: dE/dh and dE/d(pq|rs) from the converged
primal solution.

At a BPSDP optimum E = c(h, g) . x* with x* optimal for the fixed constraints
A x = b, x in the PSD cone, which do not depend on the integrals; by the
envelope theorem dE/dtheta = x* . dc/dtheta.  c is linear in (h, g) (the
packing of v2rdm_solver.cc:164-193 and tei.cc:277-314), so the gradient is the
transpose of that packing applied to x*, computed by the library's adjoint
gather kernel (v2h_batch_cost_adjoint).  No checkpointing of the iteration is
needed.  The gradient w.r.t. g is with respect to every element of the n^4
array as given (an 8-fold-unique (pq|rs) collects the sum over its images).

    E = v2rdm_energy(h, g, nalpha, nbeta, options)        # 0-d tensor
    E.backward()                                           # h.grad, g.grad

``v2rdm_energies`` solves several instances as one batch and returns one
energy per instance.  An instance that does not converge within MAXITER (or
whose state turns non-finite) returns NaN, with a RuntimeWarning: its c . x
is not the optimum and x . dc/dtheta not its gradient.
"""
from __future__ import annotations

import warnings

import numpy as np
import torch

from .batch import Batch, Problem
from .v2rdm import _POSITIVITY, resolve_options


def _problem(nmo, nalpha, nbeta, o):
    q2, g2, t1, t2 = _POSITIVITY[o["POSITIVITY"]]
    irreps = o["IRREPS"]
    nirrep = o["NIRREP"]
    if irreps is not None and nirrep is None:
        nirrep = 1 << int(np.ceil(np.log2(max(int(np.max(irreps)) + 1, 1))))
    return Problem(nmo, nalpha, nbeta, None if irreps is None else list(irreps), nirrep, 0, q2, g2,
                   bool(o["CONSTRAIN_SPIN"]), bool(o["CONSTRAIN_SZ"]), t1, t2, bool(o["CONSTRAIN_D3"]))


class _Energies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spec, *hg):
        problems, o, keep = spec
        from .batch import options as sdp_options
        opts = sdp_options(e_convergence=float(o["E_CONVERGENCE"]), r_convergence=float(o["R_CONVERGENCE"]),
                           cg_convergence=float(o["CG_CONVERGENCE"]), cg_maxiter=int(o["CG_MAXITER"]),
                           maxiter=int(o["MAXITER"]), mu_update_frequency=int(o["MU_UPDATE_FREQUENCY"]))
        B = Batch(problems, opts)
        try:
            for i, p in enumerate(problems):
                h = hg[2 * i].detach().to("cpu", torch.float64).numpy().reshape(p.nmo, p.nmo)
                g = hg[2 * i + 1].detach().to("cpu", torch.float64).numpy().reshape((p.nmo,) * 4)
                B.set_integrals(i, h, g)
            B.solve_to_convergence()
            st = [B.status(i) for i in range(B.count)]
            grads = [B.cost_adjoint(i) for i in range(B.count)]
            if keep is not None:
                keep["status"] = st
                keep["x"] = [B.x(i) for i in range(B.count)]
        finally:
            B.close()
        # The envelope gradient x* . dc/dtheta is dE/dtheta only at a converged
        # optimum: an instance that stopped at MAXITER or with a non-finite
        # state (status_bits) gets a NaN energy and NaN gradients, with a
        # warning, and leaves the other instances of the batch untouched.
        energies = []
        bad = []
        for i, s in enumerate(st):
            if s["converged"] and not s["status_bits"]:
                energies.append(s["energy_primal"])
            else:
                bad.append(i)
                energies.append(float("nan"))
                grads[i] = tuple(np.full_like(d, np.nan) for d in grads[i])
        if bad:
            warnings.warn("v2RDM SDP not converged (MAXITER reached or non-finite state) for instance(s) %s: "
                          "energy and gradient set to NaN" % bad, RuntimeWarning, stacklevel=3)
        ctx.grads = grads
        ctx.shapes = [(t.shape, t.dtype, t.device) for t in hg]
        dev = hg[0].device if hg else "cpu"
        return torch.tensor(energies, dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, gE):
        out = [None]
        gE = gE.detach().to("cpu", torch.float64)
        for i, (dh, dg) in enumerate(ctx.grads):
            for k, d in enumerate((dh, dg)):
                shape, dtype, dev = ctx.shapes[2 * i + k]
                out.append((gE[i] * torch.from_numpy(d)).reshape(shape).to(device=dev, dtype=dtype))
        return tuple(out)


def v2rdm_energies(hs, gs, electrons, options=None, info=None):
    """Energies of several fixed-orbital DQG SDPs solved as one batch.
    hs[i]: (n_i, n_i), gs[i]: (n_i,)*4 tensors; electrons[i] = (nalpha, nbeta);
    options as for v2RDMHelper, shared by every instance (IRREPS too).
    info (dict, optional) receives each instance's status and primal x."""
    o = resolve_options(options)
    problems = [_problem(int(h.shape[0]), int(na), int(nb), o) for h, (na, nb) in zip(hs, electrons)]
    return v2rdm_energies_problems(problems, hs, gs, o, info)


def v2rdm_energies_problems(problems, hs, gs, options=None, info=None):
    """As v2rdm_energies with explicit Problem descriptions (irreps etc.)."""
    o = resolve_options(options)
    hg = []
    for h, g in zip(hs, gs):
        hg += [h, g]
    return _Energies.apply((list(problems), o, info), *hg)


def v2rdm_energy(h, g, nalpha, nbeta, options=None, info=None):
    """Energy of one fixed-orbital DQG SDP as a differentiable 0-d tensor."""
    return v2rdm_energies([h], [g], [(nalpha, nbeta)], options, info)[0]
