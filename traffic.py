"""HBM traffic per kernel launch from rocprofv3 counter runs, for bench.py's
roofline.traffic.

Two counter runs of the same short bench command, one per counter:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir>/fetch -o run -- python3 bench.py --batch B ...
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir>/write -o run -- python3 bench.py --batch B ...

then

    python traffic.py --workload c2v16xB --fetch <dir>/fetch --write <dir>/write

averages FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel.  It scales
both by the calibration factor of k_copy, a plain copy of n doubles, which
moves 8 n bytes each way and runs in every bench step: factor = algorithmic
bytes / measured bytes.  The bytes per launch of every kernel (fetch + write,
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--copy-bytes", type=float, default=None,
                    help="bytes read (= bytes written) by one k_copy dispatch of the run "
                         "(default: 8 x mean n_primal of the c2v16xB workload)")
    args = ap.parse_args()
    if args.copy_bytes is None:
        import sys
        sys.path.insert(0, ROOT)
        from hilbert_amd import configs
        from hilbert_amd.structure import Structure
        nb = int(args.workload.split("x")[-1])
        nps = [Structure(p.nmo, p.nalpha, p.nbeta, p.irreps, p.nirrep).n_primal for p, _, _ in configs.c2v_16(nb)]
        args.copy_bytes = 8.0 * sum(nps) / len(nps)
    fetch = read_counter(args.fetch, "FETCH_SIZE")
    write = read_counter(args.write, "WRITE_SIZE")
    mean = lambda v: sum(v) / max(len(v), 1)
    kf = {short_name(k): mean(v) * 1024.0 for k, v in fetch.items()}
    kw = {short_name(k): mean(v) * 1024.0 for k, v in write.items()}
    cf = args.copy_bytes / kf["k_copy"] if kf.get("k_copy") else 1.0
    cw = args.copy_bytes / kw["k_copy"] if kw.get("k_copy") else 1.0
    path = os.path.join(ROOT, "profiles", "traffic.json")
    data = json.load(open(path)) if os.path.exists(path) else {}
    rec = {}
    print(f"workload {args.workload}: copy calibration fetch x{cf:.3f}, write x{cw:.3f}")
    print(f"{'kernel':24s} {'fetch MB':>10s} {'write MB':>10s} {'bytes/launch':>14s}")
    for k in sorted(set(kf) | set(kw)):
        f, w = kf.get(k, 0.0) * cf, kw.get(k, 0.0) * cw
        rec[k] = {"fetch_bytes": f, "write_bytes": w, "bytes_per_launch": f + w,
                  "fetch_factor": cf, "write_factor": cw}
        print(f"{k:24s} {f / 1e6:10.2f} {w / 1e6:10.2f} {f + w:14.4g}")
    data[args.workload] = rec
    json.dump(data, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
