"""GPU parity of the FCIDUMP writer (v2h_batch_write_fcidump) against the oracle.

Both files are read back as 24-byte records.  Labels and record order must
be identical to the oracle's restatement of fcidump.cc; the RDM values are
sums of the same primal elements in the same order, so they must match bit
for bit; the two-electron integrals come from the device's pair Gram matrix
(a different summation order over the auxiliary index than the oracle's
strided dots), so they must match to 1e-13 of the largest |value|.
"""
import numpy as np
import pytest

from hilbert_amd.batch import FCIDUMP_RECORD, Batch, Problem, options
from hilbert_amd.integrals import synthetic_df
from oracle.pyoracle import OracleProblem, fcidump_integrals, fcidump_map

pytestmark = pytest.mark.gpu

CASES = [
    # name, nfull, core, active (instance order = Pitzer), na, nb, naux, full irreps (None: C1), nirrep
    ("c1", 7, [5, 1], [0, 3, 6], 2, 1, 9, None, 1),
    ("d2h", 12, [0, 4], [1, 2, 5, 6, 8, 9], 3, 3, 14, [0, 0, 0, 1, 1, 1, 1, 2, 2, 3, 3, 3], 4),
]


def _read(path):
    rec = np.fromfile(path, dtype=FCIDUMP_RECORD)
    return rec["value"], rec["labels"]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fcidump_matches_oracle(case, tmp_path):
    name, nfull, core, active, na, nb, naux, irr, H = case
    rng = np.random.default_rng(11 + nfull)
    h, qmo, _ = synthetic_df(nfull, naux, rng, irreps=irr)
    act_irr = None if irr is None else [irr[a] for a in active]
    B = Batch([Problem(len(active), na, nb, act_irr, H if irr is not None else None)], options())
    P = OracleProblem(len(active), na, nb, act_irr if act_irr is not None else [0] * len(active), H)
    x = rng.standard_normal(P.n_primal)
    B.set_guess(0, x)
    fi, fr = str(tmp_path / "f.int"), str(tmp_path / "f.rdm")
    n_int, n_rdm = B.write_fcidump(0, fi, fr, h, qmo, core, active, enuc=7.5, full_irreps=irr)
    B.close()
    gv, gl = _read(fi)
    rv, rl = _read(fr)
    assert (len(gv), len(rv)) == (n_int, n_rdm)
    space = np.full(nfull, 2, dtype=np.int32)
    space[core] = 0
    space[active] = 1
    lab = fcidump_map(space, irr, H)
    ov, ol = fcidump_integrals(h, qmo, 7.5, lab, irr)
    assert np.array_equal(gl, ol), name
    print(name, "records", len(gv), len(rv), "max |d tei|", np.abs(gv - ov).max(), flush=True)
    assert np.abs(gv - ov).max() <= 1e-13 * max(1.0, np.abs(ov).max())
    orv, orl = P.fcidump_rdms(x, active, lab, len(core))
    assert np.array_equal(rl, orl) and np.array_equal(rv, orv), name


def test_fcidump_argument_checks(tmp_path):
    nfull = 6
    rng = np.random.default_rng(3)
    irr = [0, 0, 1, 1, 2, 3]
    h, qmo, _ = synthetic_df(nfull, 5, rng, irreps=irr)
    B = Batch([Problem(2, 1, 1, [0, 1], 4)], options())
    fi, fr = str(tmp_path / "a.int"), str(tmp_path / "a.rdm")
    with pytest.raises(Exception, match="irreps differ"):
        # full orbitals 2 and 3 are both irrep 1; the instance's orbital 0 is irrep 0
        B.write_fcidump(0, fi, fr, h, qmo, [0], [2, 3], full_irreps=irr)
    with pytest.raises(Exception, match="twice"):
        B.write_fcidump(0, fi, fr, h, qmo, [1], [1, 2], full_irreps=irr)
    with pytest.raises(Exception, match="cannot open"):
        B.write_fcidump(0, str(tmp_path / "no" / "x.int"), fr, h, qmo, [0], [1, 2], full_irreps=irr)
    B.close()
