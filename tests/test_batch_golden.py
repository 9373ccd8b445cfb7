"""The mixed-batch oracle fixture (tests/golden/batch64_d2h_oracle.json,
written by tests/golden/make_golden.py --batch) describes the instances that
the batch generator produces, and every one of them converged."""
import json
import os

from hilbert_amd import configs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "batch64_d2h_oracle.json")


def test_fixture_matches_generator():
    gold = json.load(open(GOLDEN))
    items = configs.batch64_d2h()[:gold["count"]]
    assert len(gold["instances"]) == gold["count"] == len(items)
    for (prob, _, _), ref in zip(items, gold["instances"]):
        assert ref["converged"]
        assert (prob.nmo, prob.nalpha, prob.nbeta) == (ref["nmo"], ref["nalpha"], ref["nbeta"])
        assert list(prob.irreps) == ref["irreps"]
        assert abs(ref["energy_primal"] - ref["energy_dual"]) < gold["e_convergence"]
        assert ref["eps_p"] < gold["r_convergence"] and ref["eps_d"] < gold["r_convergence"]
