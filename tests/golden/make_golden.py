"""Write tests/golden/oracle_solutions.json: converged oracle solutions of the
CPU-runnable workloads (energies, iteration counts, spin, 1-RDM and the C1
spin-free 2-RDM), used by the GPU parity tests so that they do not rerun the
oracle.  Converged to e_convergence 1e-7, r_convergence 1e-6: at the
default 1e-6 / 1e-5 two correct implementations stop anywhere inside a band
of a few 1e-6 Eh (finite-precision CG on the singular A A^T makes the
iteration trajectory implementation dependent), at the tighter thresholds
they agree to about 1e-7 Eh.

Run from the repository root:  python tests/golden/make_golden.py

With --c2v16, --big and --batch64 it writes the per-instance fixtures of the
benchmark workloads at 1e-9 / 1e-8: c2v16_oracle.json (the 128 (16e,16o) C2v
SDPs, 1-RDMs of the first 16), big16_oracle.json (the 16 (20e-24e,20o-24o)
D2h SDPs with 1-RDMs) and batch64_d2h_all_oracle.json (all 64 instances of
the mixed D2h batch).

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


def _solve_fixture(args):
    """One instance at 1e-9 / 1e-8 (MU_UPDATE_FREQUENCY 200: the optimum does
    not depend on the cadence, and the oracle gets there in about half the
    outer iterations; instance 0 of the (16e,16o) batch lands 2.4e-8 Eh from
    its cadence-1000 solve)."""
    k, prob, h, g, keep_opdm = args
    out = solve(prob, h, g, BATCH_E_CONV, BATCH_R_CONV, nthreads=2, mu_update_frequency=200)
    if not keep_opdm:
        out.pop("opdm")
    out["index"] = k
    out["irreps"] = list(prob.irreps)
    return out


def fixture_main(name, items, keep_opdm, path, workers=4):
    """Converged oracle energies (and 1-RDMs where keep_opdm(k)) of a list of
    instances, written to path after every instance so that an interrupted
    run keeps what it finished; instances already in the file are skipped."""
    from concurrent.futures import ProcessPoolExecutor, as_completed
    done = {}
    if os.path.exists(path):
        with open(path) as f:
            done = {o["index"]: o for o in json.load(f)["instances"]}
    todo = [(k, p, h, g, keep_opdm(k)) for k, (p, h, g) in enumerate(items) if k not in done]
    with ProcessPoolExecutor(max_workers=workers) as ex:
        futs = [ex.submit(_solve_fixture, t) for t in todo]
        for fu in as_completed(futs):
            o = fu.result()
            done[o["index"]] = o
            res = {"workload": name, "e_convergence": BATCH_E_CONV, "r_convergence": BATCH_R_CONV,
                   "mu_update_frequency": 200, "count": len(items),
                   "instances": [done[k] for k in sorted(done)]}
            with open(path + ".tmp", "w") as f:
                json.dump(res, f, indent=1)
            os.replace(path + ".tmp", path)
            print(f"{name}: instance {o['index']} done ({len(done)}/{len(items)})", flush=True)
    print("wrote", path)


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    if "--batch" in sys.argv:
        batch_main()
        return
    if "--c2v16" in sys.argv:
        # the (16e,16o) C2v benchmark batch: all 64 instances, 1-RDMs of 0-15
        fixture_main("c2v16x128", configs.c2v(16, 128), lambda k: k < 16, os.path.join(here, "c2v16_oracle.json"))
        return
    if "--big" in sys.argv:
        # 16 x (20e-24e, 20o-24o) D2h, naux = 3 n: all instances with 1-RDMs
        fixture_main("batch16_big", configs.batch16_big(), lambda k: True, os.path.join(here, "big16_oracle.json"))
        return
    if "--batch64" in sys.argv:
        # the remaining instances of the mixed (6e,6o)-(12e,12o) D2h batch
        fixture_main("batch64_d2h", configs.batch64_d2h(), lambda k: False,
                     os.path.join(here, "batch64_d2h_all_oracle.json"), workers=8)
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
