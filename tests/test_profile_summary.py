"""Host logic of profile_summary.py, the rocprofv3 summaries behind profiles/
and bench.py's roofline.traffic, on small synthetic rocprofv3 CSV files:
kernel names, per-dispatch counter sums, the copy calibration of FETCH_SIZE
and WRITE_SIZE, and the kernel-trace statistics."""
import argparse
import csv
import json
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
import profile_summary as ps  # noqa: E402


def _counter_csv(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for disp, kernel, value in rows:
            w.writerow(dict(Dispatch_Id=disp, Kernel_Name=kernel, Counter_Name=counter, Counter_Value=value))


def test_short_names():
    assert ps.short_name("void v2h::(anonymous namespace)::k_cg_a(v2h::BatchDev, v2h::UnitsDev)") == "cg_a"
    assert ps.short_name("k_cg_at(BatchDev, UnitsDev)") == "cg_at"
    assert ps.short_name("void k_project_tri<768, 8, 6>(BatchDev, ...)") == "project_tri192"
    assert ps.short_name("void k_project_tri<512, 4, 3>(BatchDev, ...)") == "project_tri128"
    assert ps.short_name("void k_project_tri<1024, 7, 8>(BatchDev, ...)") == "project_tri224"
    assert ps.short_name("void k_project_tri<1024, 8, 8>(BatchDev, ...)") == "project_tri256"
    assert ps.short_name("k_ref_gemm(BatchDev, RefDev, Prod, Prod, int, int)") == "k_ref_gemm"
    assert ps.short_name("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_counter_instances_are_summed_per_dispatch(tmp_path):
    d = str(tmp_path / "c")
    _counter_csv(d, "FETCH_SIZE", [(1, "k_cg_a(x)", 10.0), (1, "k_cg_a(x)", 5.0), (2, "k_cg_a(x)", 7.0),
                                   (3, "k_post(x)", 1.0)])
    got = ps.read_counter(d, "FETCH_SIZE")
    assert sorted(got["k_cg_a(x)"]) == [7.0, 15.0]
    assert got["k_post(x)"] == [1.0]


def test_traffic_applies_the_copy_calibration(tmp_path):
    mib = 2
    fetch, write = str(tmp_path / "f"), str(tmp_path / "w")
    # the copy of 2 MiB reads 2 MiB and writes 2 MiB; the fetch counter
    # reports half (factor 2), the write counter all of it (factor 1)
    _counter_csv(fetch, "FETCH_SIZE", [(1, "k_copy(double*, ...)", 1024.0),
                                       (2, "k_cg_a(x)", 100.0), (3, "k_cg_a(x)", 200.0), (4, "k_cg_a(x)", 300.0)])
    _counter_csv(write, "WRITE_SIZE", [(1, "k_copy(double*, ...)", 2048.0),
                                       (2, "k_cg_a(x)", 40.0), (3, "k_cg_a(x)", 50.0), (4, "k_cg_a(x)", 60.0)])
    out = str(tmp_path / "traffic.json")
    json.dump({"other": {"cg_a": {"bytes_per_launch": 1.0}}}, open(out, "w"))
    ps.traffic(argparse.Namespace(fetch=fetch, write=write, copy_mib=mib, workload="wl", out=out))
    data = json.load(open(out))
    assert data["other"] == {"cg_a": {"bytes_per_launch": 1.0}}  # other workloads are kept
    rec = data["wl"]["cg_a"]
    assert rec["fetch_factor"] == pytest.approx(2.0) and rec["write_factor"] == pytest.approx(1.0)
    assert rec["fetch_bytes"] == pytest.approx(200.0 * 1024 * 2.0)
    assert rec["write_bytes"] == pytest.approx(50.0 * 1024 * 1.0)
    assert rec["bytes_per_launch"] == pytest.approx(rec["fetch_bytes"] + rec["write_bytes"])
    assert data["wl"]["k_copy"]["bytes_per_launch"] == pytest.approx(2 * mib * (1 << 20))


def test_traffic_needs_the_calibration_copy(tmp_path):
    fetch, write = str(tmp_path / "f"), str(tmp_path / "w")
    _counter_csv(fetch, "FETCH_SIZE", [(1, "k_cg_a(x)", 100.0)])
    _counter_csv(write, "WRITE_SIZE", [(1, "k_cg_a(x)", 40.0)])
    with pytest.raises(SystemExit):
        ps.traffic(argparse.Namespace(fetch=fetch, write=write, copy_mib=2, workload="wl",
                                      out=str(tmp_path / "t.json")))


def test_stats_full_grid_means(tmp_path, capsys):
    d = tmp_path / "k"
    d.mkdir()
    rows = [("k_cg_a(x)", 1024, 0, 300_000), ("k_cg_a(x)", 1024, 0, 100_000), ("k_cg_a(x)", 256, 0, 20_000),
            ("k_post(x)", 512, 0, 50_000)]
    with open(d / "run_kernel_trace.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z",
                                          "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for name, grid, t0, t1 in rows:
            w.writerow(dict(Kernel_Name=name, Grid_Size_X=grid, Grid_Size_Y=1, Grid_Size_Z=1,
                            Start_Timestamp=t0, End_Timestamp=t1))
    ps.stats(str(d))
    lines = {l.split()[0]: l.split() for l in capsys.readouterr().out.splitlines()[1:]}
    calls, total_ms, _, mean_us, full_calls, full_mean_us, full_grid = lines["cg_a"][1:8]
    assert int(calls) == 3 and float(total_ms) == pytest.approx(0.42, abs=0.05)
    assert float(mean_us) == pytest.approx(140.0, abs=0.1)
    assert int(full_calls) == 2 and float(full_mean_us) == pytest.approx(200.0, abs=0.1)
    assert int(full_grid) == 1024
