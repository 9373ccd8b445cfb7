"""Benchmark: converged v2RDM SDPs per second at (16e,16o), batched on one MI355X.

A step is one batched solve, from hilbert's random guess to convergence
(E_CONVERGENCE 5e-8, R_CONVERGENCE 5e-7 by default; each outer iteration's CG
stops at a relative residual of 0.007 x max(primal, dual error), --cg-factor;
mu is rebalanced every 250 outer iterations, --mu-update-frequency),
of B independent (16e,16o) DQG SDPs
(B = 128 by default: a batch of 128 keeps the projection's latency-bound
launches busier than 64 and solves about 12 % more SDPs per second; 192 adds
3 % more) with C2v irrep labels drawn at random and synthetic molecular-like
integrals (naux = 48), every instance with its own seed.  Integrals and
tables are resident in HBM before the timed region; the timed region is
reset + solve.

Printed (one JSON line): value = converged SDPs/s of the batch, where an SDP
counts when the stopping test holds and its energy is within 1e-6 Eh of the
oracle's (per-instance fixture tests/golden/c2v16_oracle.json); the single
instance rate and the CPU oracle rate are reported beside it; `roofline` is
the dominant kernel's achieved rate: the mean work of its launches (each
launch processes the instances still iterating) over the mean duration of its
launches, timed with HIP events on the library's stream inside the timed
region.

With --workload df the line measures the density-fitted integral intake
instead (v2h_batch_set_df_integrals): (Q|mn) of a 100-orbital space with 400
auxiliary functions contracted into the (16e,16o) active-space Hamiltonian
with 10 frozen-core orbitals and packed into c.  value = intakes/s of the
device part (HIP events from the first contraction kernel to the packed c,
(Q|mn) already in HBM); the call rate including the host-to-device copy of
(Q|mn) is reported beside it.

With --workload fcidump it measures the FCIDUMP writer
(v2h_batch_write_fcidump) on the same 100-orbital space: value = integral
tables/s of its device part, the pair Gram matrix (Q|mn)^T (Q|mn) over all
5050 orbital pairs (HIP events around k_pair_gram, (Q|mn) in HBM); the rate
of whole writes (device pass, copy back, 3.2 M integral records and the RDM
records written to files) is reported beside it.

With --workload orbopt it measures one orbital-optimisation call of
v2RDM-CASSCF (v2h_batch_optimize_orbitals, ORBOPT_MAXITER 10) on the same
100-orbital space with the (16e,16o) HF-guess densities: value = calls/s of
the whole call (densities, generalized Fock, exact diagonal Hessian, steps,
exponentials, transforms of the resident (Q|mn) and the DF intakes; integrals
already in HBM).  The roofline is the FP64 rate of the call's device work
outside the intakes (HIP events) against the FP64 peak.

Usage: python bench.py [--steps K] [--warmup W] [--batch B] [--gpus 1] [--workload sdp|df|fcidump|orbopt]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix)
# E_CONVERGENCE, R_CONVERGENCE of the timed solves (--e-convergence,
# --r-convergence): 5e-8 / 5e-7 stop nearly every instance within 1e-6 Eh of
# the optimum; hilbert's defaults 1e-6 / 1e-5 stop about half of them further
# away (the band those thresholds allow is some 1e-5 Eh wide)
CONV = [5e-8, 5e-7]


def bytes_per_instance(kind, n_p, n_d):
    """Algorithmic HBM bytes one launch moves for one instance."""
    return 8 * {"cg_at": n_p + 3 * n_d,        # write q; read r, w, p
                "cg_a": n_p + 8 * n_d,         # read q; read+write y, r, p, w
                "pre": 3 * n_p + 5 * n_d,      # read c-z, x, A^T y; write 3; read b, y
                "post": 4 * n_p + n_d}[kind]   # read y, x, c; write A^T y, U


# with the refinement on, the 65..256 classes time only their fallback
# launches, so their work per launch is an upper bound
PROJ_CLASSES = (("project_jacobi16", 0, 16), ("project_tri64", 17, 64), ("project_tri128", 65, 128),
                ("project_tri192", 129, 192), ("project_tri224", 193, 224), ("project_tri256", 225, 256),
                ("project_jacobi_big", 257, 1 << 30))


def project_flops(dims):
    """Nominal FP64 flops of projecting blocks of these sizes: (16/3) d^3 each
    ((4/3) d^3 reduction + 2 d^3 back-transform + 2 d^3 for U diag(l+) U^T)."""
    return sum(16.0 / 3.0 * float(d) ** 3 for d in dims if d > 0)


MU_FREQ = [250]    # MU_UPDATE_FREQUENCY (hilbert's default 1000)
CG_FACTOR = [0.007]  # CG stopping rule: relative residual below this x max(eps_p, eps_d)


def make_batch(items, maxiter):
    from hilbert_amd.batch import Batch, options
    B = Batch([p for p, _, _ in items],
              options(e_convergence=CONV[0], r_convergence=CONV[1], maxiter=maxiter, mu_update_frequency=MU_FREQ[0]))
    B.set_cg_tolerance_factor(CG_FACTOR[0])
    for i, (_, h, g) in enumerate(items):
        B.set_integrals(i, h, g)
    return B


def reset_all(B):
    for i in range(B.count):
        B.set_guess(i, None)
        B.reset(i)


def gradients(B):
    """dE/dh and dE/d(pq|rs) of every instance from its converged x (the
    adjoint of the cost packing, v2h_batch_cost_adjoint), copied to the host."""
    return [B.cost_adjoint(i) for i in range(B.count)]


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        data = json.load(open(path))
    except (OSError, ValueError):
        return None
    rec = data.get(workload, {}).get(kernel)
    return None if rec is None else rec.get("bytes_per_launch")


def executions(kind, oiter, iiter):
    """Launches of `kind` that one instance takes part in, from its outer and
    CG iteration counts: every CG solve runs one step more than it counts
    (the step whose residual stops it)."""
    if kind in ("cg_at", "cg_a"):
        return iiter + oiter
    if kind == "pre":
        return oiter + 1
    return oiter  # post and the projection classes


def roofline(B, kt, counts, workload, dom=None, proj=None):
    """Roofline dict of kernel kind `dom` (default: the dominant kernel, the
    largest total time in the timed region).  Work per launch is the mean
    over the kernel's launches of the work of the instances each launch
    processed (instances leave a launch when their CG, or their solve, has
    stopped); time per launch is the mean HIP event time of the kernel's
    launches."""
    n_p, n_d = B.n_primal, B.n_dual
    totals = {k: v[0] for k, v in kt.items() if v[1]}
    if dom is None:
        dom = max(totals, key=totals.get)
    ms, cnt = kt[dom]
    per_launch_ms = ms / max(cnt, 1)
    traffic = load_traffic(dom, workload)
    if dom == "project_refine":
        # one "launch" = the refinement launches of one outer iteration; its
        # work = the flops of the d x d x d products it ran (2 d^3 each,
        # counted per block on the device: K, G, S and X + XE per step, the
        # final X diag(l+) X^T), over the timed region
        work = proj["flops"] / max(cnt, 1)
        achieved = work / (per_launch_ms * 1e-3) / 1e12
        return dom, {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic, "kernel": "k_ref_* (refinement)",
                     "launches": cnt, "per_launch_ms": per_launch_ms, "work_per_launch": work,
                     "work": "2 d^3 FP64 flops per d x d x d product of the refinement (device counters), "
                             "per outer iteration"}
    if dom.startswith("project"):
        lo, hi = [(a, b) for k, a, b in PROJ_CLASSES if k == dom][0]
        work = sum(executions(dom, o, it) * project_flops([d for d in B.block_dims(i) if lo <= d <= hi])
                   for i, (o, it) in enumerate(counts)) / max(cnt, 1)
        achieved = work / (per_launch_ms * 1e-3) / 1e12
        return dom, {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic, "kernel": dom,
                     "launches": cnt, "per_launch_ms": per_launch_ms, "work_per_launch": work,
                     "work": "(16/3) d^3 FP64 flops per block, mean over launches"}
    work = sum(executions(dom, o, it) * bytes_per_instance(dom, n_p[i], n_d[i])
               for i, (o, it) in enumerate(counts)) / max(cnt, 1)
    achieved = work / (per_launch_ms * 1e-3) / 1e9
    return dom, {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_" + dom,
                 "launches": cnt, "per_launch_ms": per_launch_ms, "work_per_launch": work,
                 "work": "FP64 vector bytes of the instances each launch processed (tables excluded), "
                         "mean over launches"}


def cpu_baseline(item, mean_oiter, seconds, gpu_energy):
    """Oracle BPSDP (C + OpenMP + LAPACK) on one instance, run to convergence
    when that fits in about `seconds` (estimated from a first chunk of
    iterations), else for that many seconds and converted to SDPs/s with the
    batch's mean outer-iteration count.  A converged run also checks the GPU's
    energy of the same instance."""
    from oracle.pyoracle import OracleProblem
    prob, h, g = item
    P = OracleProblem(prob.nmo, prob.nalpha, prob.nbeta, prob.irreps, prob.nirrep, t1=prob.t1, t2=prob.t2,
                      d3=prob.d3)
    c = P.pack_c(h, g)
    cores = min(os.cpu_count() or 1, 16)
    t0 = time.time()
    P.solve(c, maxiter=5, nthreads=cores, e_convergence=CONV[0], r_convergence=CONV[1])
    per_it = (time.time() - t0) / 5
    full = per_it * 1.5 * mean_oiter <= seconds
    n = 100000 if full else max(5, int(seconds / max(per_it, 1e-6)))
    t0 = time.time()
    r = P.solve(c, maxiter=n, nthreads=cores, e_convergence=CONV[0], r_convergence=CONV[1])
    dt = time.time() - t0
    out = {"unit": "SDPs/s", "cores": cores, "kind": "port"}
    if r["converged"]:
        out.update(value=1.0 / dt, seconds=dt, oiter=r["oiter"],
                   energy_oracle=r["energy_primal"], energy_gpu=gpu_energy,
                   energy_abs_diff=abs(r["energy_primal"] - gpu_energy),
                   sample=f"oracle BPSDP on instance 0 to convergence ({r['oiter']} outer iterations, "
                          f"{r['iiter']} CG steps, {dt:.1f} s, {cores} OpenMP threads, LAPACK dsyevd)")
    else:
        per_it = dt / max(r["oiter"], 1)
        out.update(value=1.0 / (per_it * mean_oiter), seconds_per_outer_iteration=per_it,
                   sample=f"oracle BPSDP on instance 0 for its first {r['oiter']} outer iterations "
                          f"({r['iiter']} CG steps, {dt:.1f} s, {cores} OpenMP threads); SDPs/s = 1 / "
                          f"(s per outer iteration x {mean_oiter:.0f} outer iterations per SDP)")
    # one thread, as the reference runs (its logs are single-threaded): the
    # first outer iterations of the same solve, scaled by the full count
    its = out.get("oiter", mean_oiter)
    k = max(5, min(40, int(10.0 / max(per_it * cores, 1e-3))))
    t0 = time.time()
    r1 = P.solve(c, maxiter=k, nthreads=1, e_convergence=CONV[0], r_convergence=CONV[1])
    per1 = (time.time() - t0) / max(r1["oiter"], 1)
    out["one_thread"] = {"value": 1.0 / (per1 * its), "unit": "SDPs/s", "cores": 1,
                         "seconds_per_outer_iteration": per1,
                         "sample": f"first {r1['oiter']} outer iterations of instance 0 on one thread, "
                                   f"scaled by {its:.0f} outer iterations per SDP"}
    return out


def energy_check(item, maxiter, gpu_energy, cpu):
    """Instance 0 solved again on the GPU at E_CONVERGENCE 1e-9, R_CONVERGENCE
    1e-8: the SDP optimum to about 1e-8 Eh.  At the bench thresholds a solver
    stops anywhere in a band around it (CG on the singular A A^T makes the
    stopping point depend on the last bits), so the GPU's and the oracle's
    energies at those thresholds are both compared with it."""
    from hilbert_amd.batch import Batch, options
    prob, h, g = item
    B = Batch([prob], options(e_convergence=1e-9, r_convergence=1e-8, maxiter=maxiter))
    B.set_integrals(0, h, g)
    t0 = time.perf_counter()
    solve_with_progress(B, "tight reference (instance 0)")
    dt = time.perf_counter() - t0
    st = B.status(0)
    B.close()
    e_ref = st["energy_primal"]
    out = {"instance": 0, "thresholds": [1e-9, 1e-8], "converged": bool(st["converged"]), "seconds": dt,
           "energy_reference": e_ref, "gpu_abs_diff": abs(gpu_energy - e_ref)}
    if cpu and "energy_oracle" in cpu:
        out["oracle_abs_diff"] = abs(cpu["energy_oracle"] - e_ref)
    return out


PROGRESS_CHUNK = [1000]


FIXTURES = {"sdp": "c2v16_oracle.json", "d2h64": "batch64_d2h_all_oracle.json", "big16": "big16_oracle.json"}


def load_fixture(workload, nmo, batch, positivity, d3):
    """Per-instance oracle energies of the benchmark batch (tests/golden/,
    written by tests/golden/make_golden.py --c2v16 / --batch64 / --big at
    1e-9 / 1e-8): {index: energy}, or {} when the workload has no fixture."""
    if positivity != "DQG" or d3 or (workload == "sdp" and nmo != 16):
        return {}
    path = os.path.join(ROOT, "tests", "golden", FIXTURES[workload])
    try:
        data = json.load(open(path))
    except (OSError, ValueError):
        return {}
    return {o["index"]: o["energy_primal"] for o in data["instances"] if o["index"] < batch and o["converged"]}


def check_energies(B, ref, tol=1e-6):
    """(instances counted as converged, max |E - E_oracle|, failing indices):
    an instance counts when the solver's stopping test holds and, where the
    oracle fixture has it, its energy is within tol of the oracle's."""
    ok, worst, bad = 0, 0.0, []
    for i in range(B.count):
        st = B.status(i)
        good = bool(st["converged"])
        if i in ref:
            d = abs(st["energy_primal"] - ref[i])
            worst = max(worst, d)
            good = good and d <= tol
        if good:
            ok += 1
        else:
            bad.append(i)
    return ok, worst, bad


def solve_with_progress(B, label, chunk=None):
    """B.solve_to_convergence() with a progress line on stderr per chunk of
    outer iterations (chunked solves give the same bits as one call)."""
    chunk = chunk or PROGRESS_CHUNK[0]
    t0 = time.perf_counter()
    while True:
        B.solve(chunk)
        st = [B.status(i) for i in range(B.count)]
        done = sum(s["converged"] for s in st)
        print(f"[bench] {label}: {max(s['oiter'] for s in st)} outer iterations, {done}/{B.count} converged, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if all(s["converged"] or s["oiter"] >= B.opts.maxiter or s["status_bits"] for s in st):
            return done


DF_NFULL, DF_NQ, DF_NCORE, DF_NACT = 100, 400, 10, 16


def df_flops(n, nq, ncore):
    """FP64 flops of one intake: (pq|rs) = Pa Pa^T over n^2 active pairs, the
    exchange M M^T, the Coulomb dots and E_fzc."""
    return 2.0 * (n ** 4 * nq + n * n * ncore * nq + n * n * nq + ncore * ncore * nq)


def bench_df(args):
    import numpy as np
    from hilbert_amd.batch import Batch, Problem, options
    from hilbert_amd.integrals import synthetic_df
    rng = np.random.default_rng(2024)
    irr = np.sort(rng.integers(0, 4, DF_NFULL)).astype(np.int32)
    h, qmo, _ = synthetic_df(DF_NFULL, DF_NQ, rng, irreps=irr)
    # core: the first DF_NCORE orbitals; active: the next DF_NACT (Pitzer order)
    core = np.arange(DF_NCORE, dtype=np.int32)
    active = np.arange(DF_NCORE, DF_NCORE + DF_NACT, dtype=np.int32)
    prob = Problem(DF_NACT, DF_NACT // 2, DF_NACT // 2, irr[active].tolist(), 4)
    B = Batch([prob], options())
    for _ in range(args.warmup):
        B.set_df_integrals(0, h, qmo, core, active)
    B.set_profiling(True)
    B.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        efzc = B.set_df_integrals(0, h, qmo, core, active)
    call_s = (time.perf_counter() - t0) / args.steps
    dev_ms, launches = B.kernel_times()["df_intake"]
    dev_ms /= max(1, launches)
    c_gpu = B.c(0)
    B.close()
    flops = df_flops(DF_NACT, DF_NQ, DF_NCORE)
    achieved = flops / (dev_ms * 1e-3) / 1e12
    cpu = None
    if args.cpu_seconds > 0:
        from oracle.pyoracle import OracleProblem, df_integrals
        P = OracleProblem(DF_NACT, DF_NACT // 2, DF_NACT // 2, irr[active].tolist(), 4)
        t0 = time.perf_counter()
        reps = 0
        while True:
            ha, ga, e_ref = df_integrals(h, qmo, core, active)
            c_ref = P.pack_c(ha, ga)
            reps += 1
            if time.perf_counter() - t0 > min(args.cpu_seconds, 10.0):
                break
        dt = (time.perf_counter() - t0) / reps
        err = float(np.abs(c_gpu - c_ref).max())
        cpu = {"value": 1.0 / dt, "unit": "intakes/s", "cores": 1, "kind": "port",
               "sample": f"{reps} full intakes (oracle orc_df_integrals + orc_pack_c, one thread); "
                         f"max |c_gpu - c_oracle| = {err:.1e}, |dE_fzc| = {abs(efzc - e_ref):.1e}"}
    out = {
        "metric": "DF integral intakes/s (device part), 100 orbitals, 400 aux, 10 core, (16e,16o) active",
        "value": 1e3 / dev_ms, "unit": "intakes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dev_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "density-fitted intake: (Q|mn) -> active (pq|rs), core Fock h', E_fzc, packed c",
                   "nfull": DF_NFULL, "naux": DF_NQ, "ncore": DF_NCORE, "nmo": DF_NACT, "nirrep": 4},
        "call_rate_with_h2d": {"value": 1.0 / call_s, "unit": "intakes/s",
                               "note": "whole API call: host-to-device copy of (Q|mn) and c norms included"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "note": "whole device part of one intake (5-6 kernels) against the FP64 peak"},
        "cpu_baseline": cpu,
        "host": host_info(),
    }
    print(json.dumps(out), flush=True)


def orbopt_flops(nfull, nq, no, evals, applies):
    """FP64 flops of the GEMMs of one orbital call: per gradient evaluation
    (O)^4 integrals, Z, the Fock product, (pp|tu) and (ps|pu); per applied
    rotation the 14 Taylor terms of exp(K) and the transforms of every (Q|mn)
    slice, h and the accumulated rotation (the squarings of exp(K), at most a
    few nfull^3 products, are not counted)."""
    n2 = no * no
    per_eval = 2.0 * (n2 * n2 * nq * 2 + no * nfull * no * nq + nfull * nq * n2 + nfull * n2 * nq)
    per_apply = 2.0 * nfull ** 3 * (14 + 2 * nq + 3)
    return evals * per_eval + applies * per_apply


def ordered_df(rng, irr):
    """Synthetic DF integrals of a molecule-like shape: (Q|mn) scaled to
    (pp|pp) ~ 0.4 Eh and h ordered core < active < virtual."""
    import numpy as np
    from hilbert_amd.integrals import synthetic_df
    h, qmo, _ = synthetic_df(DF_NFULL, DF_NQ, rng, irreps=irr)
    qmo *= 0.2
    diag = np.concatenate([-4.0 + 0.1 * np.arange(DF_NCORE), -1.2 + 0.08 * np.arange(DF_NACT),
                           0.6 + 0.05 * np.arange(DF_NFULL - DF_NCORE - DF_NACT)])
    h[np.diag_indices(DF_NFULL)] = diag
    return h, qmo


def bench_orbopt(args):
    import numpy as np
    from hilbert_amd.batch import Batch, Problem, options
    rng = np.random.default_rng(2024)
    # C2v labels in turn: the (16e,16o) active space has 4 orbitals per irrep,
    # the shape of the batched-solve workload
    irr = (np.arange(DF_NFULL) % 4).astype(np.int32)
    irr[DF_NCORE:DF_NCORE + DF_NACT] = np.repeat(np.arange(4), DF_NACT // 4)
    h, qmo = ordered_df(rng, irr)
    core = np.arange(DF_NCORE, dtype=np.int32)
    active = np.arange(DF_NCORE, DF_NCORE + DF_NACT, dtype=np.int32)
    prob = Problem(DF_NACT, DF_NACT // 2, DF_NACT // 2, irr[active].tolist(), 4)
    B = Batch([prob], options(maxiter=20000))
    # the densities of a converged SDP at the starting orbitals (untimed)
    B.set_df_integrals(0, h, qmo, core, active)
    B.set_guess(0, None)
    print("[orbopt] integrals set, solving the SDP", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    while True:
        B.solve(200)
        sdp = B.status(0)
        print(f"[orbopt] SDP at the starting orbitals: {sdp['oiter']} outer iterations, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if sdp["converged"] or sdp["oiter"] >= 20000 or sdp["status_bits"]:
            break
    for _ in range(args.warmup):
        B.set_df_integrals(0, h, qmo, core, active)
        B.optimize_orbitals(0, irr)
    B.set_profiling(True)
    B.reset_kernel_times()
    call_s = 0.0
    res = None
    for k in range(args.steps):
        B.set_df_integrals(0, h, qmo, core, active)
        t0 = time.perf_counter()
        res = B.optimize_orbitals(0, irr)
        call_s += time.perf_counter() - t0
        print(f"[orbopt] call {k}: {res['iterations']} iterations, dE {res['dE']:.3e}, |g| {res['gnorm']:.3e}",
              file=sys.stderr, flush=True)
    call_s /= args.steps
    kt = B.kernel_times()
    orb_ms = kt["orbopt"][0] / args.steps
    U_gpu, h_gpu, _ = B.df_orbitals(0)
    x = B.x(0)
    # the whole v2RDM-CASSCF loop from the starting orbitals (ORBOPT_FREQUENCY
    # 500), SDP state warm from the converged fixed-orbital solve
    casscf = None
    if args.casscf:
        B.set_profiling(False)
        e0 = B.set_df_integrals(0, h, qmo, core, active)
        e_fixed = float(B.c(0) @ B.x(0)) + e0

        def prog(m, last):
            print(f"[orbopt] CASSCF macro iteration {m}: E {last[0]['energy']:.10f}, |g| {last[0]['gnorm']:.2e}, "
                  f"dE {last[0]['dE']:.2e}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        out = B.solve_casscf(orbopt_frequency=500, full_irreps=irr, progress=prog)[0]
        wall = time.perf_counter() - t0
        casscf = {"seconds": wall, "energy": out["energy"], "fixed_orbital_energy": e_fixed,
                  "macro_iterations": out["macro"], "sdp_outer_iterations": out["oiter"],
                  "converged": bool(out["sdp_converged"] and out["orbopt_converged"]),
                  "note": "one instance, warm start from the converged fixed-orbital SDP; "
                          "energies c.x + E_fzc without nuclear repulsion"}
    B.close()
    flops = orbopt_flops(DF_NFULL, DF_NQ, DF_NCORE + DF_NACT, res["evaluations"], res["rotations"])
    achieved = flops / (orb_ms * 1e-3) / 1e12
    cpu = None
    if args.cpu_seconds > 0:
        from oracle import orbopt as oo
        from oracle.pyoracle import OracleProblem
        P = OracleProblem(DF_NACT, DF_NACT // 2, DF_NACT // 2, irr[active].tolist(), 4)
        D1, d = oo.spatial_rdms(P, x)
        D, dd = oo.occupied_rdms(D1, d, DF_NCORE)
        pz = oo.pitzer_order(irr[active])
        occ = list(core) + [int(active[u]) for u in pz]
        pairs = oo.rotation_pairs(DF_NFULL, core, active, irr)
        Bd = oo.unpack_df(qmo, DF_NFULL)
        print("[orbopt] oracle call", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        h2, _, Uo, info = oo.optimize_orbitals(h, Bd, D, dd, occ, pairs)
        dt = time.perf_counter() - t0
        try:
            from threadpoolctl import threadpool_info
            threads = max([t.get("num_threads", 1) for t in threadpool_info()] or [1])
        except Exception:
            threads = None
        cpu = {"value": 1.0 / dt, "unit": "orbital calls/s", "cores": threads, "kind": "port",
               "sample": f"one call of the oracle's optimize_orbitals (numpy, BLAS threads as reported) on the "
                         f"same densities: {info['iterations']} iterations; max |U_gpu - U_oracle| = "
                         f"{float(np.abs(U_gpu - Uo).max()):.1e}, |E_gpu - E_oracle| = "
                         f"{abs(res['energy'] - info['energy']):.1e}"}
    out = {
        "metric": "v2RDM-CASSCF orbital-optimisation calls/s (ORBOPT_MAXITER 10), 100 orbitals, 400 aux, "
                  "10 core, (16e,16o) active",
        "value": 1.0 / call_s, "unit": "calls/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": call_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "orbital step: densities, generalized Fock, exact diagonal Hessian, steps, "
                               "exp(K), (Q|mn) transforms, DF intakes", "nfull": DF_NFULL, "naux": DF_NQ,
                   "ncore": DF_NCORE, "nmo": DF_NACT, "nirrep": 4, "rotations": res["nrot"],
                   "densities": "converged DQG SDP at the starting orbitals (%d outer iterations)" % sdp["oiter"],
                   "iterations": res["iterations"], "gradient_evaluations": res["evaluations"],
                   "rotations_applied": res["rotations"], "dE": res["dE"], "gnorm": res["gnorm"]},
        "device_ms": {"orbopt_kernels": orb_ms, "df_intakes": kt["df_intake"][0] / args.steps},
        "casscf": casscf,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "note": "GEMM flops of the call (transforms of 400 slices dominate) over the HIP-event "
                             "time of its device work outside the intakes (k_dgemm_mfma on the FP64 matrix cores, k_dgemm_small)"},
        "cpu_baseline": cpu,
        "host": host_info(),
    }
    print(json.dumps(out), flush=True)


def bench_fcidump(args):
    import tempfile
    import numpy as np
    from hilbert_amd.batch import Batch, Problem, options
    from hilbert_amd.integrals import synthetic_df
    rng = np.random.default_rng(2024)
    irr = np.sort(rng.integers(0, 4, DF_NFULL)).astype(np.int32)
    h, qmo, _ = synthetic_df(DF_NFULL, DF_NQ, rng, irreps=irr)
    core = np.arange(DF_NCORE, dtype=np.int32)
    active = np.arange(DF_NCORE, DF_NCORE + DF_NACT, dtype=np.int32)
    prob = Problem(DF_NACT, DF_NACT // 2, DF_NACT // 2, irr[active].tolist(), 4)
    B = Batch([prob], options())
    B.set_guess(0, None)
    ntri = DF_NFULL * (DF_NFULL + 1) // 2
    with tempfile.TemporaryDirectory() as d:
        fi, fr = os.path.join(d, "bench.int"), os.path.join(d, "bench.rdm")
        for _ in range(args.warmup):
            B.write_fcidump(0, fi, fr, h, qmo, core, active, enuc=1.0, full_irreps=irr)
        B.set_profiling(True)
        B.reset_kernel_times()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            n_int, n_rdm = B.write_fcidump(0, fi, fr, h, qmo, core, active, enuc=1.0, full_irreps=irr)
        call_s = (time.perf_counter() - t0) / args.steps
        x = B.x(0)
    dev_ms, launches = B.kernel_times()["pair_gram"]
    dev_ms /= max(1, launches)
    B.close()
    flops = 2.0 * DF_NQ * ntri * (ntri + 1) / 2       # every pair product a <= b
    achieved = flops / (dev_ms * 1e-3) / 1e12
    cpu = None
    if args.cpu_seconds > 0:
        from oracle.pyoracle import OracleProblem, fcidump_integrals, fcidump_map
        space = np.full(DF_NFULL, 2, dtype=np.int32)
        space[core] = 0
        space[active] = 1
        lab = fcidump_map(space, irr, 4)
        P = OracleProblem(DF_NACT, DF_NACT // 2, DF_NACT // 2, irr[active].tolist(), 4)
        t0 = time.perf_counter()
        fcidump_integrals(h, qmo, 1.0, lab, irr)
        P.fcidump_rdms(x, active, lab, DF_NCORE)
        dt = time.perf_counter() - t0
        cpu = {"value": 1.0 / dt, "unit": "FCIDUMP record sets/s", "cores": 1, "kind": "port",
               "sample": "one full set of integral and RDM records in memory (oracle orc_fcidump_*, one thread; "
                         "the oracle counts the records in a first pass and fills them in a second)"}
    out = {
        "metric": "FCIDUMP integral tables/s (device pair Gram), 100 orbitals, 400 aux",
        "value": 1e3 / dev_ms, "unit": "tables/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dev_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "FCIDUMP writer: (pq|rs) for all 5050 x 5050 pairs from (Q|mn), records of "
                               "integrals and of the (16e,16o) RDMs written to files",
                   "nfull": DF_NFULL, "naux": DF_NQ, "ncore": DF_NCORE, "nmo": DF_NACT, "nirrep": 4,
                   "integral_records": n_int, "rdm_records": n_rdm},
        "whole_write": {"value": 1.0 / call_s, "unit": "writes/s",
                        "note": "copy-in of (Q|mn), device Gram, 204 MB copy back, record loops, file writes"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "note": "k_pair_gram, algorithmic flops of the pair products a <= b (FP64 FMA on VALU; "
                             "MI355X's FP64 vector and matrix peaks are equal)"},
        "cpu_baseline": cpu,
        "host": host_info(),
    }
    print(json.dumps(out), flush=True)


def host_info():
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    return {"nproc": os.cpu_count(), "cpu": model}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 1; df: 200 intakes)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--maxiter", type=int, default=20000)
    ap.add_argument("--single", type=int, default=4,
                    help="also solve the first K instances one at a time, each alone on the GPU (0 = skip)")
    ap.add_argument("--cpu-seconds", type=float, default=60.0,
                    help="CPU oracle budget: a full solve of instance 0 if it fits, else a sample (0 = skip)")
    ap.add_argument("--energy-check", type=int, default=1,
                    help="solve instance 0 again at 1e-9 / 1e-8 and compare the energies with it (0 = skip)")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="after the solve, three plain copies of this many MiB (k_copy), which calibrate "
                         "the traffic counters of a profile (0 = none)")
    ap.add_argument("--nmo", type=int, default=16,
                    help="sdp workload: orbitals (= electrons) of each instance, C2v labels at random")
    ap.add_argument("--positivity", choices=("DQG", "DQGT1", "DQGT2", "DQGT1T2"), default="DQG",
                    help="sdp workload: the N-representability conditions")
    ap.add_argument("--d3", type=int, default=0, help="sdp workload: add the D3 condition (CONSTRAIN_D3)")
    ap.add_argument("--chunk", type=int, default=1000, help="outer iterations between progress lines")
    ap.add_argument("--energies", default="", help="sdp workload: write the per-instance status of step 0 here")
    ap.add_argument("--autotune", type=int, default=300,
                    help="sdp workloads: tune the launch configuration after this many outer iterations of the "
                         "first warm-up solve (0: keep the defaults)")
    ap.add_argument("--e-convergence", type=float, default=5e-8, help="E_CONVERGENCE of the timed solves")
    ap.add_argument("--r-convergence", type=float, default=5e-7, help="R_CONVERGENCE of the timed solves")
    ap.add_argument("--cg-factor", type=float, default=0.007,
                    help="CG stopping rule of the batched and single solves: relative residual below this x the "
                         "larger of the primal and dual errors (the library's default, the oracle's, is 0.003)")
    ap.add_argument("--mu-update-frequency", type=int, default=250,
                    help="MU_UPDATE_FREQUENCY of the batched and single solves: outer iterations between mu "
                         "rebalancings (hilbert's default 1000)")
    ap.add_argument("--casscf", type=int, default=1,
                    help="orbopt workload: also run the whole v2RDM-CASSCF loop once (0 = skip)")
    ap.add_argument("--workload", choices=("sdp", "d2h64", "big16", "df", "fcidump", "orbopt"), default="sdp",
                    help="sdp: the batched (16e,16o) solve; d2h64: the 64 (6e,6o)-(12e,12o) D2h SDPs; big16: "
                         "the 16 (20e,20o)-(24e,24o) D2h SDPs with their adjoint gradients; df: the "
                         "density-fitted integral intake; fcidump: the FCIDUMP writer; orbopt: one v2RDM-CASSCF "
                         "orbital step")
    ap.add_argument("--refine-steps", type=int, default=5,
                    help="refinement steps before a block falls back to the tridiagonal eigensolver")
    ap.add_argument("--projection", choices=("refined", "tridiagonal"), default="refined",
                    help="PSD blocks of 65..256 rows: warm-started refinement on the FP64 matrix cores with the "
                         "tridiagonal fallback, or the tridiagonal eigensolver alone")
    args = ap.parse_args()
    PROGRESS_CHUNK[0] = max(1, args.chunk)
    MU_FREQ[0] = args.mu_update_frequency
    CG_FACTOR[0] = args.cg_factor
    CONV[0], CONV[1] = args.e_convergence, args.r_convergence
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: batches of independent SDPs shard across GPUs without "
              "communication (one process per GPU, each with its share of the batch); this benchmark "
              "measures one GPU.", file=sys.stderr)
        sys.exit(2)

    from hilbert_amd import _native, configs
    if _native.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.workload == "df":
        if args.steps is None:
            args.steps = 200
        bench_df(args)
        return
    if args.workload == "fcidump":
        if args.steps is None:
            args.steps = 5
        bench_fcidump(args)
        return
    if args.workload == "orbopt":
        if args.steps is None:
            args.steps = 5
        bench_orbopt(args)
        return
    if args.steps is None:
        args.steps = 1

    if args.workload == "d2h64":
        args.batch = min(args.batch, 64)  # the mixed D2h workload has 64 instances
        items = configs.batch64_d2h(args.batch)
    elif args.workload == "big16":
        args.batch = min(args.batch, 16)
        items = configs.batch16_big(args.batch)
    else:
        items = configs.c2v(args.nmo, args.batch)
    t1 = args.positivity in ("DQGT1", "DQGT1T2")
    t2 = args.positivity in ("DQGT2", "DQGT1T2")
    for prob, _, _ in items:
        prob.t1, prob.t2, prob.d3 = t1, t2, bool(args.d3)
    B = make_batch(items, args.maxiter)
    B.set_projection(B.PROJECT_REFINED if args.projection == "refined" else B.PROJECT_TRIDIAGONAL)
    B.set_refine_params(max_steps=args.refine_steps)
    B.set_profiling(True)
    grads = args.workload == "big16"
    tuned = None
    for w in range(args.warmup):
        reset_all(B)
        if w == 0 and args.autotune:
            # the tuner times its candidates on the live warm-up solve (its
            # knobs change no result bit), after the first outer iterations;
            # with the kernel timers off, since with them on the graph modes
            # launch one by one
            from hilbert_amd import autotune
            B.solve(args.autotune)
            B.set_profiling(False)
            tuned = autotune.tune(B, log=lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
            B.set_profiling(True)
        solve_with_progress(B, f"warmup {w}")
        if grads:
            gradients(B)
    B.reset_kernel_times()
    proj0 = B.projection_stats()

    ref = load_fixture(args.workload, args.nmo, args.batch, args.positivity, args.d3)
    step_times, conv, conv_solver, oiters, counts = [], [], [], [], [[0, 0] for _ in range(args.batch)]
    worst, failing = 0.0, set()
    for k in range(args.steps):
        reset_all(B)
        t0 = time.perf_counter()
        nconv = solve_with_progress(B, f"step {k}")
        if grads:
            gradients(B)
        step_times.append(time.perf_counter() - t0)
        conv_solver.append(nconv)
        ok, w, bad = check_energies(B, ref)
        conv.append(ok)
        worst = max(worst, w)
        failing.update(bad)
        if k == 0 and args.energies:
            with open(args.energies, "w") as f:
                json.dump([{"index": i, **B.status(i)} for i in range(B.count)], f)
        oiters.append([B.status(i)["oiter"] for i in range(B.count)])
        for i in range(B.count):
            st = B.status(i)
            counts[i][0] += st["oiter"]
            counts[i][1] += st["iiter"]
    kt = B.kernel_times()
    total = sum(step_times)
    value = sum(conv) / total
    mean_oiter = sum(sum(o) for o in oiters) / max(1, sum(len(o) for o in oiters))
    st = [B.status(i) for i in range(B.count)]
    n = args.nmo
    cond = args.positivity + ("+D3" if args.d3 else "")
    workload = {"sdp": f"c2v{n}x{args.batch}", "d2h64": f"d2h{args.batch}",
                "big16": f"big{args.batch}"}[args.workload] + ("" if cond == "DQG" else "_" + cond.lower().replace("+", ""))
    proj1 = B.projection_stats()
    proj = {"mode": args.projection, "refine_steps": args.refine_steps,
            "flops": proj1["flops"] - proj0["flops"],
            "refined": proj1["refined"] - proj0["refined"],
            "fallback": proj1["fallback"] - proj0["fallback"]}
    dom, roof = roofline(B, kt, counts, workload, proj=proj)
    # the HBM-bound gather kernel of the CG step, whichever kernel dominates
    roof_cg = roofline(B, kt, counts, workload, "cg_a")[1] if dom != "cg_a" else None
    roof_ref = (roofline(B, kt, counts, workload, "project_refine", proj=proj)[1]
                if dom != "project_refine" and kt.get("project_refine", (0, 0))[1] else None)
    energy0 = B.status(0)["energy_primal"]
    B.close()

    if args.calibration_copy_mib > 0:
        n = args.calibration_copy_mib * (1 << 20) // 8
        ms = _native.calibration_copy(n)
        print(f"[bench] calibration copy of {args.calibration_copy_mib} MiB: {ms:.3f} ms, "
              f"{2 * 8 * n / ms / 1e6:.0f} GB/s", file=sys.stderr, flush=True)

    single = None
    if args.single:
        # one instance at a time: the single-SDP drop-in use; the outer
        # iteration count of one instance varies a lot from instance to
        # instance, so the rate is taken over several
        secs, oits, good = [], [], 0
        for k in range(min(args.single, len(items))):
            S = make_batch(items[k:k + 1], args.maxiter)
            S.set_projection(S.PROJECT_REFINED if args.projection == "refined" else S.PROJECT_TRIDIAGONAL)
            S.set_refine_params(max_steps=args.refine_steps)
            reset_all(S)
            t0 = time.perf_counter()
            solve_with_progress(S, f"single instance {k}")
            secs.append(time.perf_counter() - t0)
            st0 = S.status(0)
            oits.append(st0["oiter"])
            good += int(st0["converged"] and (k not in ref or abs(st0["energy_primal"] - ref[k]) <= 1e-6))
            S.close()
        single = {"value": good / sum(secs), "unit": "SDPs/s", "instances": len(secs),
                  "within_1e-6_of_oracle": good, "seconds": secs, "oiter": oits}

    cpu = None
    if args.cpu_seconds > 0:
        print(f"[bench] CPU oracle on instance 0 (budget {args.cpu_seconds:.0f} s)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(items[0], mean_oiter, args.cpu_seconds, energy0)
    echeck = None
    if args.energy_check:
        echeck = energy_check(items[0], 200000, energy0, cpu)

    if args.workload == "sdp":
        metric = f"converged v2RDM SDPs/s (ΔE≤1e-6 Eh), batched vs single GPU vs CPU, ({n}e,{n}o)"
        wl = (f"{args.batch} x ({n}e,{n}o) {cond} SDPs, C2v labels at random, naux {3 * n}, "
              "one batched solve to convergence per step")
        cfg = {"nmo": n, "electrons": n, "nirrep": 4}
    else:
        ns = [p.nmo for p, _, _ in items]
        lo, hi = min(ns), max(ns)
        metric = (f"converged v2RDM SDPs/s (ΔE≤1e-6 Eh), {args.batch} x ({lo}e,{lo}o)-({hi}e,{hi}o) D2h"
                  + (" with adjoint gradients" if grads else ""))
        wl = (f"{args.batch} x ({lo}e,{lo}o)-({hi}e,{hi}o) {cond} SDPs, D2h labels at random, "
              f"naux {'3 n' if grads else 30}, one batched solve to convergence per step"
              + (" and dE/dh, dE/d(pq|rs) of every instance" if grads else ""))
        cfg = {"nmo": [lo, hi], "electrons": [lo, hi], "nirrep": 8}
    out = {
        "metric": metric,
        "value": value, "unit": "SDPs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "batch": args.batch, **cfg, "projection": args.projection,
                   "e_convergence": CONV[0], "r_convergence": CONV[1], "positivity": args.positivity,
                   "mu_update_frequency": MU_FREQ[0], "cg_tolerance_factor": CG_FACTOR[0],
                   "constrain_d3": bool(args.d3), "constrain_spin": True, "constrain_sz": True},
        "converged": sum(conv), "attempted": args.batch * args.steps,
        "converged_check": {"stopping_test": sum(conv_solver), "within_1e-6_of_oracle": sum(conv),
                            "instances_with_oracle_energy": len(ref), "max_abs_dE_vs_oracle": worst,
                            "failing_instances": sorted(failing),
                            "oracle": f"tests/golden/{FIXTURES[args.workload]}: oracle BPSDP at 1e-9 / 1e-8 "
                                      "per instance"},
        "projection": proj,
        "tuning": tuned,
        "mean_outer_iterations": mean_oiter,
        "mean_cg_per_outer": sum(s["iiter"] for s in st) / max(1, sum(s["oiter"] for s in st)),
        "single_gpu": single,
        "batched_over_single": (value / single["value"]) if single else None,
        "cpu_baseline": cpu,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
        "energy_check": echeck,
        "roofline": roof,
        "roofline_cg_a": roof_cg,
        "roofline_refine": roof_ref,
        "kernel_ms": {k: round(v[0], 3) for k, v in kt.items() if v[1]},
        "kernel_launches": {k: v[1] for k, v in kt.items() if v[1]},
        "host": host_info(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
