"""Option registry of the v2RDMHelper mirror (plugin.cc names and defaults);
host only."""
import pytest

from hilbert_amd.v2rdm import DEFAULTS, resolve_options


def test_defaults_follow_plugin():
    o = resolve_options()
    assert o["E_CONVERGENCE"] == 1e-6 and o["R_CONVERGENCE"] == 1e-5 and o["MAXITER"] == 50
    assert o["CG_CONVERGENCE"] == 1e-9 and o["CG_MAXITER"] == 10000 and o["MU_UPDATE_FREQUENCY"] == 1000
    assert o["POSITIVITY"] == "DQG" and o["CONSTRAIN_SPIN"] and o["CONSTRAIN_SZ"]
    assert o["TPDM_GUESS"] == "RANDOM" and o["SDP_SOLVER"] == "BPSDP"


def test_case_insensitive_names_and_values():
    o = resolve_options({"positivity": "dq", "maxiter": 20000})
    assert o["POSITIVITY"] == "DQ" and o["MAXITER"] == 20000


@pytest.mark.parametrize("opt", [{"POSITIVITY": "DQGT3"}, {"SDP_SOLVER": "RRSDP"},
                                 {"CONSTRAIN_E3": True}, {"GPC_CONSTRAINTS": "1RDM"}, {"TAU_PARAMETER": 1.5}])
def test_unimplemented_settings_raise(opt):
    with pytest.raises(NotImplementedError):
        resolve_options(opt)


def test_unknown_option_raises():
    with pytest.raises(KeyError):
        resolve_options({"NOT_AN_OPTION": 1})
    assert "IRREPS" in DEFAULTS


def test_guess_values():
    assert resolve_options({"TPDM_GUESS": "hf"})["TPDM_GUESS"] == "HF"
    with pytest.raises(ValueError):
        resolve_options({"TPDM_GUESS": "SAD"})
