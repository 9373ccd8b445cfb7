"""Summaries of rocprofv3 runs of bench.py, for profiles/.

stats -- from a kernel-trace run

    rocprofv3 --kernel-trace --stats --output-format csv -d <dir> -o run -- python3 bench.py ...
    python profile_summary.py stats <dir>

prints, per kernel, the dispatch count and mean duration over all dispatches
and over the dispatches with the kernel's largest grid.  The largest grid is
a launch over the whole batch (instances that converged drop out of later
launches), which is what bench.py's per-kernel roofline times.

counters -- from a counter run of any command

    rocprofv3 --pmc C1 C2 ... --output-format csv -d <dir> -o run -- python3 bench.py ...
    python profile_summary.py counters <dir> --counter C1 --counter C2 ...

prints, per kernel, the dispatch count and each counter's mean per dispatch.

traffic -- from two counter runs of the same short bench command

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir>/fetch -o run -- python3 bench.py --batch B \
        --calibration-copy-mib M ...
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir>/write -o run -- python3 bench.py --batch B \
        --calibration-copy-mib M ...
    python profile_summary.py traffic --workload c2v16xB --copy-mib M --fetch <dir>/fetch --write <dir>/write

(bench.py with --calibration-copy-mib M) averages FETCH_SIZE and WRITE_SIZE
(KiB per dispatch) per kernel (summed over the counter's instances within a
dispatch).  It scales both by the calibration factor of
the largest k_copy dispatch, bench's plain copy of M MiB, which moves M MiB
each way: factor = M MiB / measured bytes.  The bytes per launch of every kernel (fetch + write,
scaled) go into profiles/traffic.json under the workload name, and a text
summary goes to stdout.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
from collections import defaultdict

ROOT = os.path.dirname(os.path.abspath(__file__))


def short_name(kernel: str) -> str:
    m = re.search(r"(k_[A-Za-z0-9_]+)(<([^>]*)>)?", kernel)
    if not m:
        return kernel
    name = m.group(1)
    if name == "k_project_tri" and m.group(3):
        nt = m.group(3).split(",")[0].strip()
        ri = m.group(3).split(",")[1].strip() if m.group(3).count(",") >= 1 else ""
        if nt == "1024":
            return "project_tri224" if ri == "7" else "project_tri256"
        return {"256": "project_tri64", "512": "project_tri128", "768": "project_tri192"}.get(nt, name)
    if name == "k_project_wave":
        return "project_tri64"
    return {"k_cg_at": "cg_at", "k_cg_a": "cg_a", "k_pre": "pre", "k_post": "post"}.get(name, name)


def read_counter(d: str, counter: str):
    """{kernel: [values per dispatch]} from the counter_collection CSV(s) of a run."""
    out = defaultdict(list)
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    for f in files:
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                key = row.get("Kernel_Name", "")
                disp = row.get("Dispatch_Id") or row.get("Correlation_Id") or str(len(out[key]))
                out[key].append((disp, float(row["Counter_Value"])))
    res = {}
    for k, v in out.items():
        per = defaultdict(float)
        for disp, val in v:  # sum over the counter's instances within a dispatch
            per[disp] += val
        res[k] = list(per.values())
    return res


def stats(d: str):
    files = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no kernel_trace.csv under {d}")
    per = defaultdict(list)  # kernel -> [(grid, ns)]
    for f in files:
        with open(f) as fh:
            for row in csv.DictReader(fh):
                grid = 1
                for ax in ("X", "Y", "Z"):
                    grid *= int(row.get(f"Grid_Size_{ax}") or 1)
                ns = int(row["End_Timestamp"]) - int(row["Start_Timestamp"])
                per[short_name(row["Kernel_Name"])].append((grid, ns))
    total_all = sum(ns for v in per.values() for _, ns in v)
    print(f"{'kernel':22s} {'calls':>8s} {'total ms':>10s} {'%':>6s} {'mean us':>9s} "
          f"{'full calls':>10s} {'full mean us':>12s} {'full grid':>10s}")
    for k, v in sorted(per.items(), key=lambda kv: -sum(ns for _, ns in kv[1])):
        tot = sum(ns for _, ns in v)
        gmax = max(g for g, _ in v)
        full = [ns for g, ns in v if g == gmax]
        print(f"{k:22s} {len(v):8d} {tot / 1e6:10.1f} {100.0 * tot / total_all:6.1f} {tot / len(v) / 1e3:9.1f} "
              f"{len(full):10d} {sum(full) / len(full) / 1e3:12.1f} {gmax:10d}")


def counters(args):
    """Per kernel: dispatches and the mean of each counter per dispatch."""
    vals = {c: read_counter(args.dir, c) for c in args.counter}
    kernels = sorted({short_name(k) for v in vals.values() for k in v})
    print(f"{'kernel':24s} {'dispatches':>10s} " + " ".join(f"{c:>28s}" for c in args.counter))
    for k in kernels:
        row, n = [], 0
        for c in args.counter:
            v = [x for kk, xs in vals[c].items() if short_name(kk) == k for x in xs]
            n = max(n, len(v))
            row.append(f"{(sum(v) / len(v)) if v else float('nan'):28.6g}")
        print(f"{k:24s} {n:10d} " + " ".join(row))


def traffic(args):
    fetch = read_counter(args.fetch, "FETCH_SIZE")
    write = read_counter(args.write, "WRITE_SIZE")
    mean = lambda v: sum(v) / max(len(v), 1)
    kf = {short_name(k): mean(v) * 1024.0 for k, v in fetch.items()}
    kw = {short_name(k): mean(v) * 1024.0 for k, v in write.items()}
    copy_bytes = args.copy_mib * float(1 << 20)
    cfk = max((max(v) for k, v in fetch.items() if short_name(k) == "k_copy"), default=0.0) * 1024.0
    cwk = max((max(v) for k, v in write.items() if short_name(k) == "k_copy"), default=0.0) * 1024.0
    if cfk <= 0.0 or cwk <= 0.0:
        raise SystemExit("no k_copy dispatch in the counter runs (bench.py --calibration-copy-mib)")
    cf, cw = copy_bytes / cfk, copy_bytes / cwk
    path = getattr(args, "out", None) or os.path.join(ROOT, "profiles", "traffic.json")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    data = json.load(open(path)) if os.path.exists(path) else {}
    rec = {}
    print(f"workload {args.workload}: calibration copy of {args.copy_mib} MiB measured fetch {cfk / 1e6:.1f} MB, "
          f"write {cwk / 1e6:.1f} MB -> factors fetch x{cf:.3f}, write x{cw:.3f}")
    print(f"{'kernel':24s} {'dispatches':>10s} {'fetch MB':>10s} {'write MB':>10s} {'bytes/launch':>14s}")
    for k in sorted(set(kf) | set(kw)):
        f, w = kf.get(k, 0.0) * cf, kw.get(k, 0.0) * cw
        n = max(len(v) for kk, v in fetch.items() if short_name(kk) == k) if k in kf else 0
        rec[k] = {"fetch_bytes": f, "write_bytes": w, "bytes_per_launch": f + w,
                  "fetch_factor": cf, "write_factor": cw}
        print(f"{k:24s} {n:10d} {f / 1e6:10.2f} {w / 1e6:10.2f} {f + w:14.4g}")
    data[args.workload] = rec
    json.dump(data, open(path, "w"), indent=1, sort_keys=True)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--counter", action="append", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--workload", required=True)
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-mib", type=int, required=True,
                   help="size of bench.py's calibration copy (--calibration-copy-mib)")
    t.add_argument("--out", default=None, help="JSON file to update (default profiles/traffic.json)")
    args = ap.parse_args()
    if args.cmd == "stats":
        stats(args.dir)
    elif args.cmd == "counters":
        counters(args)
    else:
        traffic(args)


if __name__ == "__main__":
    main()
