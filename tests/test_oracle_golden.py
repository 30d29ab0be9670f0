"""Pins the CPU oracle against the reference's own known-answer tests.

Each vector below is the value a ticit unit test asserts (file:line cited);
the oracle is a restatement, so these are what make it trustworthy.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ticit_amd import plan_text

GOLDEN = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1


def test_splitmix64_published_vectors():
    # src/random.rs:89-103
    v, s = O.splitmix_next(0)
    assert v == 0xE220A8397B1DCDAF
    v, s = O.splitmix_next(s)
    assert v == 0x6E789E6AA1B965F4
    v, s = O.splitmix_next(s)
    assert v == 0x06C45D188009454F
    assert s == (GOLDEN * 3) & MASK
    assert O.splitmix_next(1)[0] == 0x910A2DEC89025CC1
    assert O.splitmix_next(42)[0] == 0xBDD732262FEB6E95


def test_rand_float_pinned():
    # src/random.rs:105-116
    u, _ = O.rand_float(0)
    assert u == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert u == 0.88331080821364261


def test_block_seed_definition():
    # src/random.rs:151-162
    assert O.block_seed(2, 3, 4) == 2 ^ ((GOLDEN * 5) & MASK) ^ ((0xBF58476D1CE4E5B9 * 4) & MASK)
    a = O.block_seed(0x7EED0000, 0, 0)
    assert a != O.block_seed(0x7EED0000, 1, 0)
    assert a != O.block_seed(0x7EED0000, 0, 1)
    assert a != O.block_seed(0x5EED1234, 0, 0)


def test_categorical_walk_is_inclusive_and_falls_through():
    # src/random.rs:139-148
    assert O.categorical_row(0, [0.0, 0.0])[0] == 1
    assert O.categorical_row(0, [1.0, 0.0])[0] == 0


def test_certain_bernoulli_consumes_no_randomness():
    # src/random.rs:118-129
    b, s = O.bernoulli(7, 0.0)
    assert (b, s) == (False, 7)
    b, s = O.bernoulli(7, 1.0)
    assert (b, s) == (True, 7)
    _, s = O.bernoulli(7, 0.5)
    assert s != 7
    with pytest.raises(O.OracleError):
        O.bernoulli(1, -0.1)
    with pytest.raises(O.OracleError):
        O.bernoulli(1, float("nan"))


def test_insert_zero_bit():
    # src/active.rs:395-400
    assert O.insert_zero_bit(0b1011, 0) == 0b10110
    assert O.insert_zero_bit(0b1011, 2) == 0b10011
    assert O.insert_zero_bit(0b1011, 4) == 0b01011


def test_default_batch_count():
    # src/sampler/batch/mod.rs:712-717
    assert O.default_batch(10) == 32
    assert O.default_batch(0) == 2048
    assert O.default_batch(15) == 1


def test_exact_rows_of_deterministic_circuit():
    # src/sampler/prepared.rs:1207-1245
    blob = plan_text("M !0\nM 1\nDETECTOR rec[-2] rec[-1]\n"
                     "OBSERVABLE_INCLUDE(0) rec[-2]\nOBSERVABLE_INCLUDE(2) rec[-1]\n"
                     "EXP_VAL Z0\n")
    r = O.sample(blob, 3, 7, bit_packed=False)
    assert r["measurements"].reshape(-1).tolist() == [1, 0, 1, 0, 1, 0]
    assert r["detectors"].reshape(-1).tolist() == [1, 1, 1]
    assert r["observables"].reshape(-1).tolist() == [1, 0, 0, 1, 0, 0, 1, 0, 0]
    assert r["observable_ones"].tolist() == [3, 0, 0]
    assert r["exp_vals"].reshape(-1).tolist() == [1.0] * 3
    p = O.sample(blob, 3, 7, bit_packed=True)
    assert p["measurements"].reshape(-1).tolist() == [1, 1, 1]
    assert p["detectors"].reshape(-1).tolist() == [1, 1, 1]
    assert p["observables"].reshape(-1).tolist() == [1, 1, 1]
    assert p["exp_vals"].tolist() == r["exp_vals"].tolist()


def test_expectation_probes_exact_and_non_destructive():
    # src/sampler/batch/mod.rs:745-766: Z0 of |0> is 1, X0 is 0 (a dormant
    # probe); after H T the X, Y, Z probes read 1/sqrt2, 1/sqrt2, 0; the final
    # measurement still reads 0 (T_DAG H undo the state)
    blob = plan_text("EXP_VAL Z0 X0\nH 0\nT 0\nEXP_VAL X0 Y0 Z0\nT_DAG 0\nH 0\nM 0\n")
    r = O.sample(blob, 5, 71, chunk_shots=0, batch_size=0, bit_packed=False)
    want = [1.0, 0.0, 2 ** -0.5, 2 ** -0.5, 0.0]
    assert r["exp_vals"].shape == (5, 5)
    assert np.abs(r["exp_vals"] - np.array(want)[None, :]).max() <= 1e-12
    assert r["measurements"].reshape(-1).tolist() == [0] * 5


def test_packed_rows_little_bit_order():
    # src/sampler/prepared.rs:1199-1205; ticit_py/tests/test_api.py:73-88
    blob = plan_text("X 0 1 2 3 4 5 6 7 10\nM 0 1 2 3 4 5 6 7 8 9 10\n")
    r = O.sample(blob, 2, 1, bit_packed=True)
    assert r["measurements"].tolist() == [[0xFF, 0x04], [0xFF, 0x04]]


def test_t_gate_window():
    # src/sampler/batch/mod.rs:719-731: H T M, 200 shots, seed 23, batch 32
    blob = plan_text("H 0\nT 0\nM 0\n")
    r = O.sample(blob, 200, 23, batch_size=32, bit_packed=False)
    ones = int(r["measurements"].sum())
    assert 50 < ones < 150


def test_seed_determinism_and_sensitivity():
    # src/sampler/batch/mod.rs:733-743
    blob = plan_text("X_ERROR(0.25) 0\nH 1\nT 1\nH 1\nM 0 1\n")
    a = O.sample(blob, 300, 99)
    b = O.sample(blob, 300, 99)
    c = O.sample(blob, 300, 100)
    assert np.array_equal(a["measurements"], b["measurements"])
    assert not np.array_equal(a["measurements"], c["measurements"])


def test_thread_count_invariance_of_counts():
    # src/sampler/prepared.rs:1340-1375: 1 vs 3 workers give equal counts
    from ticit_amd import workloads as w
    blob = plan_text(w.surface_code_t(3, 3, 1e-2, 1, 0))
    a = O.sample(blob, 10000, 5, threads=1)
    b = O.sample(blob, 10000, 5, threads=3)
    assert a["observable_ones"].tolist() == b["observable_ones"].tolist()
    assert sorted(map(bytes, a["measurements"])) == sorted(map(bytes, b["measurements"]))


def test_per_shot_contract_is_reference_at_unit_chunks():
    """Per-shot streams == reference sampling with one-shot chunks/blocks,
    and a shot's row does not depend on which other shots run with it."""
    from ticit_amd import workloads as w
    blob = plan_text(w.surface_code_t(3, 2, 1e-2, 1, 0))
    full = O.sample_per_shot(blob, 300, 9)
    part = O.sample(blob, 300, 9, chunk_shots=1, batch_size=1, threads=1)
    assert np.array_equal(full["measurements"], part["measurements"])


def test_presample_sparse_hit_rate():
    # gpu/plan.rs:1568-1577 window: p = 1/8 Bernoulli group over 20000 draws
    blob = plan_text("X_ERROR(0.125) 0\nM 0\n")
    planes = O.presample(blob, 20000, 1234)
    hits = int(np.unpackbits(planes.view(np.uint8), bitorder="little").sum())
    assert 2300 <= hits <= 2700


def test_postselection_returns_only_survivor_rows():
    # src/sampler/prepared.rs:1248-1273 (batch 16: no branch is drawn)
    blob = plan_text("EXP_VAL Z1\nM !1\nDETECTOR rec[-1]\nX_ERROR(0.5) 0\nM 0\nDETECTOR rec[-1]\n"
                     "OBSERVABLE_INCLUDE(0) rec[-2]\n", postselection_mask=[0, 1])
    r = O.sample(blob, 128, 9, batch_size=16, bit_packed=False)
    rows = int(r["counts"][2])
    assert 0 < rows < 128 and int(r["counts"][1]) == 128 - rows
    assert r["measurements"].tolist() == [[1, 0]] * rows
    assert r["detectors"].tolist() == [[1, 0]] * rows
    assert r["observables"].reshape(-1).tolist() == [1] * rows
    assert r["observable_ones"].tolist() == [rows]
    assert r["exp_vals"].reshape(-1).tolist() == [1.0] * rows


def test_source_discards_and_the_passed_mask_are_unioned():
    # src/sampler/prepared.rs:1276-1307
    text = "M !0\nDETECTOR rec[-1]\nM 0\nDISCARD rec[-1]\n"
    c = O.sample(plan_text(text), 8, 1, bit_packed=False)["counts"]
    assert (int(c[2]), int(c[1])) == (8, 0)
    c = O.sample(plan_text(text, postselection_mask=[1]), 8, 1, bit_packed=False)["counts"]
    assert (int(c[2]), int(c[1])) == (0, 8)
    c = O.sample(plan_text("M !0\nDISCARD rec[-1]\n"), 8, 1, bit_packed=False)["counts"]
    assert (int(c[2]), int(c[1])) == (0, 8)


def test_an_always_fired_postselected_detector_discards_every_shot():
    # src/sampler/prepared.rs:1111-1130: estimate_logical_error_rate
    # postselects every detector (prepared.rs:1037)
    r = O.sample(plan_text("M !0\nOBSERVABLE_INCLUDE(0) rec[-1]\n", postselection_mask=[]), 5, 1)
    assert (int(r["counts"][1]), int(r["counts"][2]), int(r["counts"][3])) == (0, 5, 5)
    r = O.sample(plan_text("M !0\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-1]\n", postselection_mask=[1]), 5, 1)
    assert (int(r["counts"][1]), int(r["counts"][2]), int(r["counts"][3])) == (5, 0, 0)
    assert r["measurements"].shape[0] == 0


def test_postselected_rows_are_the_unpostselected_rows_of_the_survivors():
    """At one shot per stream, discarding a shot changes no other shot's
    draws, so the survivors' rows are the plain run's rows of those shots."""
    text = ("H 0\nT 0\nH 0\nX_ERROR(0.2) 1\nM 0 1\nDETECTOR rec[-1]\nH 2\nT 2\nM 2\n"
            "DETECTOR rec[-1] rec[-2]\nOBSERVABLE_INCLUDE(0) rec[-3]\n")
    plain = O.sample_per_shot(plan_text(text), 500, 3, bit_packed=False)
    post = O.sample_per_shot(plan_text(text, postselection_mask=[1, 0]), 500, 3, bit_packed=False)
    keep = plain["detectors"][:, 0] == 0
    assert int(post["counts"][2]) == int(keep.sum()) and int(post["counts"][1]) == 500 - int(keep.sum())
    assert np.array_equal(post["measurements"], plain["measurements"][keep])
    assert np.array_equal(post["detectors"], plain["detectors"][keep])
    assert int(post["observable_ones"][0]) == int(plain["observables"][keep, 0].sum())
