"""Host logic of the autotuner (hilbert_amd/autotune.py): candidate
configurations and the batch signature, without a GPU (the batch is a
stand-in with block dimensions)."""
import json

import numpy as np

from hilbert_amd import autotune

BASE = {"class_lane": [0, 0, 2, 3, 1, 1, 2], "fallback_lane": [0, 0, 2, 3, 3, 2, 2], "refine_lane": 1,
        "cg_margin": 2, "xcd_remap": 0, "graphs": 0, "pipeline": 0}


class _FakeBatch:
    def __init__(self, dims, mode=1):
        self.dims, self.count, self.mode = dims, len(dims), mode

    def block_dims(self, i):
        return np.array(self.dims[i], dtype=np.int32)

    def projection_mode(self):
        return self.mode


def test_lane_presets_are_valid_tunings():
    for name, (cl, fl, rl) in autotune.LANE_PRESETS.items():
        assert len(cl) == 7 and len(fl) == 7, name
        assert all(0 <= v <= 3 for v in cl + fl), name
        assert 1 <= rl <= 3, name


def test_candidates_change_one_knob_and_leave_the_base_alone():
    base = json.loads(json.dumps(BASE))
    c = autotune._with(base, "lanes", "serial")
    assert c["class_lane"] == [0] * 7 and c["fallback_lane"] == [0] * 7 and c["refine_lane"] == 1
    assert {k: v for k, v in c.items() if k not in ("class_lane", "fallback_lane", "refine_lane")} == \
        {k: v for k, v in BASE.items() if k not in ("class_lane", "fallback_lane", "refine_lane")}
    assert autotune._with(base, "cg_margin", 8)["cg_margin"] == 8
    assert base == BASE
    assert set(autotune.CG_MARGINS) <= set(range(65))


def test_signature_depends_on_the_block_multiset_and_the_mode():
    a = _FakeBatch([[66, 16, 4], [100, 32]])
    b = _FakeBatch([[100, 32], [66, 16, 4]])  # same instances, other order
    c = _FakeBatch([[66, 16, 4], [100, 33]])
    d = _FakeBatch([[66, 16, 4], [100, 32]], mode=0)
    assert autotune.signature(a) == autotune.signature(b)
    assert autotune.signature(a) != autotune.signature(c)
    assert autotune.signature(a) != autotune.signature(d)


def test_apply_cached_reads_the_signature_entry(tmp_path):
    path = str(tmp_path / "t.json")
    a = _FakeBatch([[66, 16, 4]])
    got = {}
    a.set_tuning = lambda cfg: got.update(cfg)
    assert autotune.apply_cached(a, path) is None
    json.dump({autotune.signature(a): {"tuning": BASE}}, open(path, "w"))
    assert autotune.apply_cached(a, path) == BASE and got == BASE
