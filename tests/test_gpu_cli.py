"""The command line mirror on the device: printed counts and record files
equal the library's sample() and the oracle's rows (gpu-marked)."""
import numpy as np
import pytest

from oracle import oracle as O
from ticit_amd import __main__ as cli
from ticit_amd import plan_text

pytestmark = pytest.mark.gpu

TEXT = ("H 0\nT 0\nH 0\nX_ERROR(0.1) 0 1\nM 0 1\nDETECTOR rec[-1] rec[-2]\nH 2\nM 2\nDETECTOR rec[-1]\n"
        "X 3\nDEPOLARIZE1(0.05) 3\nM 3\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-4]\n"
        "EXP_VAL Z0\n")


def printed(out: str) -> dict:
    return dict(line.split(" ", 1) for line in out.strip().splitlines())


def test_records_out_writes_the_oracle_rows(hip, tmp_path, capsys):
    path = tmp_path / "c.ticit"
    path.write_text(TEXT)
    prefix = str(tmp_path / "rows")
    shots, seed = 3000, 4
    assert cli.main([str(path), "-n", str(shots), "--seed", str(seed), "--records-out", prefix,
                     "--postselection-mask", "1,0,0"]) == 0
    p = printed(capsys.readouterr().out)
    blob = plan_text(TEXT, [1, 0, 0])
    o = O.sample_per_shot(blob, shots, seed, bit_packed=False)
    accepted = int(o["counts"][2])
    assert (int(p["shots"]), int(p["discarded"]), int(p["accepted"]), int(p["logical_errors"])) == \
        tuple(int(v) for v in o["counts"])
    assert p["keep_records"] == "true" and p["discard_rate"] == cli.rate((shots - accepted) / shots)
    meas = np.fromfile(prefix + ".measurements.u8", dtype=np.uint8).reshape(accepted, -1)
    det = np.fromfile(prefix + ".detectors.u8", dtype=np.uint8).reshape(accepted, -1)
    obs = np.fromfile(prefix + ".observables.u8", dtype=np.uint8).reshape(accepted, -1)
    assert np.array_equal(meas, o["measurements"]) and np.array_equal(det, o["detectors"])
    assert np.array_equal(obs, o["observables"])
    ev = np.fromfile(prefix + ".exp_vals.f64", dtype="<f8")
    assert ev.shape == (accepted,)
    meta = open(prefix + ".meta").read()
    assert meta.startswith(f"rows {accepted}\nmeasurements 4\ndetectors 3\nobservables 1\nexpectations 1\n")


def test_count_only_and_normalized_runs(hip, tmp_path, capsys):
    path = tmp_path / "c.ticit"
    path.write_text(TEXT)
    assert cli.main([str(path), "-n", "2000", "--count-only", "--normalize-syndromes",
                     "--postselection-mask", "1,1,1"]) == 0
    p = printed(capsys.readouterr().out)
    assert p["keep_records"] == "false" and p["tier"] == "lane"
    assert int(p["accepted"]) + int(p["discarded"]) == 2000 and 0 < int(p["accepted"]) < 2000
