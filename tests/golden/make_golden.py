"""Write tests/golden/oracle_solutions.json: converged oracle solutions of the
CPU-runnable workloads (energies, iteration counts, spin, 1-RDM and the C1
spin-free 2-RDM), used by the GPU parity tests so that they do not rerun the
oracle.  Converged to e_convergence 1e-7, r_convergence 1e-6: at the
default 1e-6 / 1e-5 two correct implementations stop anywhere inside a band
of a few 1e-6 Eh (finite-precision CG on the singular A A^T makes the
iteration trajectory implementation dependent), at the tighter thresholds
they agree to about 1e-7 Eh.

Run from the repository root:  python tests/golden/make_golden.py

With --batch it writes tests/golden/batch64_d2h_oracle.json instead: energies
and <S^2> of the first 16 instances of the mixed (6e,6o)-(12e,12o) D2h batch,
converged to e_convergence 1e-9, r_convergence 1e-8.  At 1e-7 / 1e-6 a
solver can stop while the energy is still moving by a few 1e-6 Eh (instance
11: the oracle stops 4.0e-6 Eh above its 1e-9 / 1e-8 energy), so the batch
is compared at the tighter pair.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from hilbert_amd import configs  # noqa: E402
from oracle.pyoracle import OracleProblem  # noqa: E402

E_CONV, R_CONV = 1e-7, 1e-6


def solve(prob, h, g, e_conv=E_CONV, r_conv=R_CONV, **kw):
    P = OracleProblem(prob.nmo, prob.nalpha, prob.nbeta, prob.irreps, prob.nirrep)
    t0 = time.time()
    r = P.solve(P.pack_c(h, g), maxiter=200000, e_convergence=e_conv, r_convergence=r_conv, **kw)
    out = {"nmo": prob.nmo, "nalpha": prob.nalpha, "nbeta": prob.nbeta,
           "converged": r["converged"], "energy_primal": r["energy_primal"], "energy_dual": r["energy_dual"],
           "eps_p": r["eps_p"], "eps_d": r["eps_d"], "oiter": r["oiter"], "iiter": r["iiter"],
           "spin_squared": P.spin_squared(r["x"]), "opdm": P.opdm(r["x"]).ravel().tolist(),
           "seconds": time.time() - t0}
    if P.nirrep == 1:
        out["tpdm"] = P.tpdm(r["x"]).ravel().tolist()
    print(f"  E = {r['energy_primal']:.10f}  oiter {r['oiter']}  iiter {r['iiter']}  {out['seconds']:.1f} s",
          flush=True)
    return out


def batch_main(count=16):
    from concurrent.futures import ProcessPoolExecutor
    items = configs.batch64_d2h()[:count]
    with ProcessPoolExecutor(max_workers=4) as ex:
        outs = list(ex.map(_solve_small, items))
    res = {"e_convergence": BATCH_E_CONV, "r_convergence": BATCH_R_CONV, "count": count, "instances": outs}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch64_d2h_oracle.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print("wrote", path)


BATCH_E_CONV, BATCH_R_CONV = 1e-9, 1e-8


def _solve_small(item):
    prob, h, g = item
    out = solve(prob, h, g, BATCH_E_CONV, BATCH_R_CONV, nthreads=2)
    out.pop("opdm")
    out["irreps"] = list(prob.irreps)
    return out


def main():
    if "--batch" in sys.argv:
        batch_main()
        return
    res = {"e_convergence": E_CONV, "r_convergence": R_CONV}
    for name in ("hubbard8", "c1_10"):
        print(name, flush=True)
        prob, h, g = configs.CONFIGS[name]()[0]
        res[name] = solve(prob, h, g)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_solutions.json")
    with open(path, "w") as f:
        json.dump(res, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
