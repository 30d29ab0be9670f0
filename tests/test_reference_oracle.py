"""Expected detector / observable bits and the noiseless reference sample in
the CPU oracle (oracle/ref_cpu.cpp), and the host-side option checks of the
HIP library (no device needed: they fail before any device call).

Pinned by:
  * the reference's own tests: noiseless_reference_normalizes_outputs_before_counting
    (src/sampler/prepared.rs:1164-1194) and
    test_normalize_syndromes_uses_cpu_reference_sample / the "cannot be
    combined" check (ticit_py/tests/test_api.py:56-70,104-109);
  * the stream the reference shot draws from, recomputed here from the
    published SplitMix64 (Steele, Lea and Flood 2014; the constants of
    Vigna's splitmix64.c) started at state 1, with every noise symbol 0:
    coin k is bit 0 of draw k, an active branch compares draw k's top 53
    bits / 2^53 with P(true) of the closed-form one-qubit state;
  * invariants: normalized rows equal raw rows XOR the expected bits,
    flipped observables count shots - raw ones, postselection keeps exactly
    the shots whose raw postselected detectors equal the expected bits, and
    a noiseless circuit whose detectors never vary samples the reference
    in every shot (the restated CCZ fixture check, prepared.rs:1078-1108,
    on the cultivation fixture).
"""
import math
import re

import numpy as np
import pytest

import ticit_amd
from oracle import oracle as O
from ticit_amd import plan_text
from ticit_amd import workloads as w

GOLDEN = "X 0\nM 0\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0)\nOBSERVABLE_INCLUDE(1) rec[-1]\n"
MASK64 = (1 << 64) - 1


@pytest.fixture(autouse=True)
def clear_state():
    yield
    O.set_expected(None, None)
    O.set_pins(None)


def splitmix_draws(state: int, n: int) -> list:
    out = []
    for _ in range(n):
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        out.append(z ^ (z >> 31))
    return out


def unpack(rows, ncols):
    return np.unpackbits(rows, axis=1, bitorder="little")[:, :ncols]


def test_reference_sample_of_the_reference_fixture():
    d, o = O.reference_sample(plan_text(GOLDEN))
    assert d.tolist() == [1]
    assert o.tolist() == [0, 1]


def test_normalized_counts_of_the_reference_fixture():
    blob = plan_text(GOLDEN, [1])
    raw = O.sample(blob, 4, 7, bit_packed=False, observable=1)
    assert raw["counts"].tolist() == [4, 4, 0, 0]  # the detector always fires
    O.set_expected(*O.reference_sample(blob))
    r = O.sample(blob, 4, 7, bit_packed=False, observable=1)
    assert r["counts"].tolist() == [4, 0, 4, 0]  # shots, discarded, accepted, logical
    assert r["detectors"].tolist() == [[0]] * 4
    assert r["observables"].tolist() == [[0, 0]] * 4
    assert r["observable_ones"].tolist() == [0, 0]


P_TRUE = math.sin(math.pi / 8) ** 2  # P(1) of H T H |0>


@pytest.mark.parametrize("kinds", ["ccc", "mmmm", "nmc", "mnc", "cnmc", "ncmn"])
def test_reference_branches_follow_the_state_one_stream(kinds):
    """c = a fresh |+> measured (a dormant coin: bit 0 of its draw); n = an
    active measurement of H T H |0>, whose branch is true when the draw's
    uniform is below P(true) = sin^2(pi/8); m = the same with an X before the
    measurement, a frame flip that complements the record of an identically
    drawn branch. Every shot also measures a qubit behind X_ERROR(0.5),
    which the noiseless reference reads as 0."""
    lines = []
    for q, k in enumerate(kinds):
        lines.append(f"X_ERROR(0.5) {q + 10}\nM {q + 10}\nDETECTOR rec[-1]")
        lines.append({"c": f"H {q}\nM {q}", "m": f"H {q}\nT {q}\nH {q}\nX {q}\nM {q}",
                      "n": f"H {q}\nT {q}\nH {q}\nM {q}"}[k])
        lines.append("DETECTOR rec[-1]")
    blob = plan_text("\n".join(lines) + "\n")
    inf = O.info(blob)
    assert inf["ncoin"] == kinds.count("c") and inf["nmeasure"] == len(kinds) - kinds.count("c")
    want = []
    for k, v in zip(kinds, splitmix_draws(1, len(kinds))):
        branch = int(v & 1) if k == "c" else int((v >> 11) * 2.0 ** -53 < P_TRUE)
        want += [0, branch ^ int(k == "m")]
    d, _ = O.reference_sample(blob)
    assert d.tolist() == want


def test_reference_obeys_pins():
    text = "H 0\nM 0\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-1]\n"
    blob = plan_text(text)
    first = int(splitmix_draws(1, 1)[0] & 1)
    assert [a.tolist() for a in O.reference_sample(blob)] == [[first], [first]]
    for value in (False, True):
        O.set_pins(blob, [([0], value)])
        assert [a.tolist() for a in O.reference_sample(blob)] == [[int(value)], [int(value)]]


def noisy_cases():
    return {
        "surface_d3_t": w.surface_code_t(3, 3, 2e-2, 1, 0),
        "mixed": "\n".join(
            ["H 0\nT 0\nH 0\nX_ERROR(0.1) 0 1\nM 0 1\nDETECTOR rec[-1] rec[-2]\nH 2\nM 2\nDETECTOR rec[-1]",
             "X 3\nDEPOLARIZE1(0.05) 3\nM 3\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-4]",
             "OBSERVABLE_INCLUDE(2) rec[-2] rec[-1]"]) + "\n",
    }


@pytest.mark.parametrize("name", sorted(noisy_cases()))
def test_expected_bits_invert_rows_and_counts(name):
    blob = plan_text(noisy_cases()[name])
    inf = O.info(blob)
    shots = 500
    raw = O.sample_per_shot(blob, shots, 5, bit_packed=False)
    rng = np.random.default_rng(3)
    ed = rng.integers(0, 2, inf["ndetectors"]).tolist()
    eo = rng.integers(0, 2, inf["nobservables"]).tolist()
    O.set_expected(ed, eo)
    r = O.sample_per_shot(blob, shots, 5, bit_packed=False)
    assert np.array_equal(r["measurements"], raw["measurements"])
    assert np.array_equal(r["detectors"], raw["detectors"] ^ np.array(ed, dtype=np.uint8))
    assert np.array_equal(r["observables"], raw["observables"] ^ np.array(eo, dtype=np.uint8))
    want = [shots - int(n) if e else int(n) for n, e in zip(raw["observable_ones"], eo)]
    assert r["observable_ones"].tolist() == want
    assert int(r["counts"][3]) == want[0]


def test_normalized_postselection_keeps_shots_agreeing_with_the_reference():
    text = noisy_cases()["mixed"]
    mask = [1, 1, 0]
    plain, post = plan_text(text), plan_text(text, mask)
    shots = 800
    raw = O.sample_per_shot(plain, shots, 9, bit_packed=False)
    ref_d, ref_o = O.reference_sample(post)
    assert np.array_equal(O.reference_sample(plain)[0], ref_d)  # postselection plays no part
    O.set_expected(ref_d, ref_o)
    r = O.sample_per_shot(post, shots, 9, bit_packed=False)
    keep = np.all(raw["detectors"][:, :2] == ref_d[:2], axis=1)
    assert int(r["counts"][2]) == int(keep.sum()) and 0 < keep.sum() < shots
    assert np.array_equal(r["measurements"], raw["measurements"][keep])
    assert np.array_equal(r["detectors"], raw["detectors"][keep] ^ ref_d)


def zero_noise(text: str) -> str:
    return re.sub(r"\b(DEPOLARIZE1|DEPOLARIZE2|X_ERROR|Y_ERROR|Z_ERROR|M|MX|MY|MR|MRX)\([^)]*\)", r"\1(0)", text)


def test_noiseless_cultivation_samples_its_reference_in_every_shot():
    text = open(__file__.replace("test_reference_oracle.py", "golden/circuits/msc_d3_inject_cultivate_p1e-3.stim")).read()
    quiet = plan_text(zero_noise(text))
    d, o = O.reference_sample(quiet)
    r = O.sample_per_shot(quiet, 64, 4)
    inf = O.info(quiet)
    assert np.all(unpack(r["detectors"], inf["ndetectors"]) == d)
    assert np.all(unpack(r["observables"], inf["nobservables"]) == o)
    # every detector postselected, normalized: nothing is discarded
    mask = [1] * inf["ndetectors"]
    blob = plan_text(zero_noise(text), mask)
    O.set_expected(*O.reference_sample(blob))
    c = O.sample_per_shot(blob, 64, 4)["counts"]
    assert int(c[1]) == 0 and int(c[2]) == 64


def test_oracle_rejects_expected_bits_of_the_wrong_length():
    blob = plan_text(GOLDEN)
    O.set_expected([1, 0], None)
    with pytest.raises(O.OracleError, match="expected_detectors has length 2, expected 1"):
        O.sample(blob, 4, 1)
    O.set_expected(None, [1])
    with pytest.raises(O.OracleError, match="expected_observables has length 1, expected 2"):
        O.sample(blob, 4, 1)


def test_library_rejects_normalize_with_explicit_bits():
    """Host-side checks of ticit_hip_program_create_ex (they run before any
    device call); InvalidInput surfaces as a ValueError, as in ticit_py."""
    c = ticit_amd.Circuit("M 0\nDETECTOR rec[-1]\n")
    with pytest.raises(ValueError, match="cannot be combined"):
        c.compile(expected_detectors=[0], normalize_syndromes=True)
    with pytest.raises(ValueError, match="cannot be combined"):
        c.compile(expected_observables=[1], normalize_syndromes=True)
    with pytest.raises(ValueError, match="expected_detectors has length 2, expected 1"):
        c.compile(expected_detectors=[0, 1])
    with pytest.raises(ValueError, match="expected_observables has length 1, expected 0"):
        c.compile(expected_observables=[1])
