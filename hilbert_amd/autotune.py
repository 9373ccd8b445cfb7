"""Autotuner of the batched solve's launch configuration (v2h_tuning).

The knobs change no result bit (the same kernels on the same data in the
same dependency order), so the tuner can time its candidates on the live
solve: every timed window is a stretch of real outer iterations, and a tuned
solve returns the same bits as an untuned one.

Knobs, searched one at a time (coordinate descent, in this order), each
candidate timed over `window` outer iterations in `rounds` interleaved
rounds (median):

* ``graphs``: eager launches, hipGraph replay of the first CG chunk of each
  outer iteration, or of that chunk and of k_post with the whole projection;
* ``xcd_remap``: gather workgroups dealt to the 8 XCDs in contiguous ranges
  of one instance (L2 locality) or in launch order;
* ``cg_margin``: the first CG chunk of an outer iteration is the largest CG
  step count of the previous one plus this margin (larger: fewer host polls,
  more skipped launches);
* ``lanes``: which of the four streams each PSD size class, the refinement
  and its tridiagonal fallback run on.

The choice is kept per batch signature (the instances' block dimensions and
the projection mode) in a JSON cache; ``apply_cached`` sets it on a new batch
with the same signature.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

# (class_lane, fallback_lane, refine_lane) presets; classes: Jacobi <= 16,
# tridiagonal <= 64 / 128 / 192 / 224 / 256, Jacobi > 256
LANE_PRESETS = {
    "balanced": ([0, 0, 2, 3, 1, 1, 2], [0, 0, 2, 3, 3, 2, 2], 1),
    "small_on_main": ([0, 0, 1, 2, 3, 3, 1], [0, 0, 2, 3, 2, 3, 3], 1),
    "two_fallback_lanes": ([0, 1, 2, 3, 2, 3, 1], [0, 0, 2, 2, 3, 3, 3], 1),
    "refine_on_3": ([0, 0, 1, 2, 2, 1, 1], [0, 0, 1, 2, 2, 1, 1], 3),
    "serial": ([0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0, 0], 1),
}
CG_MARGINS = (0, 1, 2, 4, 8)


def default_cache() -> str:
    return os.environ.get("V2H_TUNING_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "hilbert_amd", "tuning.json"))


def signature(B) -> str:
    """Key of a batch: the sorted multiset of its instances' block dimensions
    and the projection mode."""
    dims = sorted(tuple(int(d) for d in B.block_dims(i)) for i in range(B.count))
    blob = json.dumps([dims, int(B.projection_mode())])
    return hashlib.sha1(blob.encode()).hexdigest()[:16]


def _with(cfg: dict, knob: str, value) -> dict:
    out = json.loads(json.dumps(cfg))
    if knob == "lanes":
        out["class_lane"], out["fallback_lane"], out["refine_lane"] = LANE_PRESETS[value]
        out["class_lane"], out["fallback_lane"] = list(out["class_lane"]), list(out["fallback_lane"])
    else:
        out[knob] = value
    return out


def _window(B, iters: int) -> float:
    """Seconds per outer iteration over `iters` iterations of the live solve
    (one untimed iteration first: a new configuration captures its graphs)."""
    B.solve(1)
    o0 = max(B.status(i)["oiter"] for i in range(B.count))
    t0 = time.perf_counter()
    B.solve(iters)
    dt = time.perf_counter() - t0
    o1 = max(B.status(i)["oiter"] for i in range(B.count))
    return dt / max(o1 - o0, 1)


def _running(B) -> bool:
    return any(not s["converged"] and s["status_bits"] == 0 and s["oiter"] < B.opts.maxiter
               for s in (B.status(i) for i in range(B.count)))


def tune(B, window: int = 20, rounds: int = 3, cache: str | None = None, log=None) -> dict:
    """Time the candidates on the live solve of batch B (best after its first
    few hundred outer iterations, when the per-iteration work has settled,
    and with the kernel timers off: with them on, the graph modes launch one
    by one),
    set the fastest configuration, store it under B's signature and return
    {"tuning": cfg, "ms_per_iteration": {knob: {value: ms}}, "signature"}.
    Stops early, keeping the best so far, if the batch converges."""
    best = B.tuning()
    report = {}
    for knob, values in (("graphs", (0, 1, 2)), ("xcd_remap", (1, 0)), ("cg_margin", CG_MARGINS),
                         ("lanes", tuple(LANE_PRESETS))):
        cands = [_with(best, knob, v) for v in values]
        times = [[] for _ in cands]
        for _ in range(rounds):
            for k, c in enumerate(cands):
                if not _running(B):
                    break
                B.set_tuning(c)
                times[k].append(_window(B, window))
        med = [float(np.median(t)) if t else float("inf") for t in times]
        k = int(np.argmin(med))
        if np.isfinite(med[k]):
            best = cands[k]
        report[knob] = {str(v): (round(1e3 * m, 4) if np.isfinite(m) else None) for v, m in zip(values, med)}
        if log:
            log(f"autotune {knob}: ms per outer iteration {report[knob]} -> {values[k]}")
    B.set_tuning(best)
    out = {"tuning": best, "ms_per_iteration": report, "signature": signature(B)}
    path = cache or default_cache()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = json.load(open(path)) if os.path.exists(path) else {}
        data[out["signature"]] = {"tuning": best, "ms_per_iteration": report}
        with open(path, "w") as f:
            json.dump(data, f, indent=1)
    except OSError:
        pass  # a read-only cache location only loses the memo
    return out


def apply_cached(B, cache: str | None = None) -> dict | None:
    """Set the cached configuration of B's signature, if there is one."""
    path = cache or default_cache()
    try:
        data = json.load(open(path))
    except (OSError, ValueError):
        return None
    rec = data.get(signature(B))
    if rec is None:
        return None
    B.set_tuning(rec["tuning"])
    return rec["tuning"]
