"""Host logic of bench.py's roofline line on a stand-in batch: the dominant
kernel, its algorithmic bytes per launch from the instances' iteration
counts, achieved GB/s and the fraction of the HBM peak, the refinement's
TFLOP/s from the device flop counters, and the traffic lookup."""
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
import bench  # noqa: E402


class _FakeBatch:
    def __init__(self, n_p, n_d, dims):
        self.n_primal, self.n_dual, self._dims = n_p, n_d, dims

    def block_dims(self, i):
        return self._dims[i]


def test_cg_a_roofline_by_hand():
    B = _FakeBatch([100, 120], [200, 240], [[4], [4]])
    counts = [(10, 90), (20, 180)]  # (outer, CG) iterations of each instance
    # cg_a launches an instance takes part in: CG steps + one per outer iteration
    runs = [90 + 10, 180 + 20]
    launches = 200                  # the kernel's launches (the longer instance's)
    kt = {"cg_a": (2.0, launches), "cg_at": (1.0, launches), "post": (0.5, 20)}
    dom, r = bench.roofline(B, kt, counts, "no_such_workload")
    assert dom == "cg_a"
    work = (runs[0] * 8 * (100 + 8 * 200) + runs[1] * 8 * (120 + 8 * 240)) / launches
    assert r["work_per_launch"] == pytest.approx(work)
    assert r["per_launch_ms"] == pytest.approx(2.0 / launches)
    assert r["achieved"] == pytest.approx(work / (2.0 / launches * 1e-3) / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["unit"] == "GB/s" and r["bound"] == "hbm" and r["traffic"] is None


def test_named_kernel_and_bytes_per_instance():
    B = _FakeBatch([100], [200], [[4]])
    kt = {"cg_a": (2.0, 10), "cg_at": (1.0, 10)}
    dom, r = bench.roofline(B, kt, [(1, 9)], "none", "cg_at")
    assert dom == "cg_at"
    assert r["work_per_launch"] == pytest.approx(10 * 8 * (100 + 3 * 200) / 10)
    assert bench.bytes_per_instance("post", 10, 3) == 8 * (4 * 10 + 3)
    assert bench.executions("pre", 5, 40) == 6 and bench.executions("post", 5, 40) == 5


def test_refinement_roofline_from_device_flops():
    B = _FakeBatch([100], [200], [[80]])
    kt = {"project_refine": (4.0, 2), "cg_a": (1.0, 10)}
    flops = 3.0e10
    dom, r = bench.roofline(B, kt, [(2, 20)], "none", proj={"flops": flops})
    assert dom == "project_refine"
    assert r["work_per_launch"] == pytest.approx(flops / 2)
    assert r["achieved"] == pytest.approx(flops / 2 / (2.0e-3) / 1e12)
    assert r["frac"] == pytest.approx(r["achieved"] / 78.6) and r["unit"] == "TFLOP/s"


def test_traffic_lookup_reads_the_committed_summary():
    got = bench.load_traffic("cg_a", "c2v16x128")
    assert got is not None and 5e8 < got < 5e9
    assert bench.load_traffic("cg_a", "no_such_workload") is None
