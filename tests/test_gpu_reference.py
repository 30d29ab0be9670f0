"""Expected detector / observable bits and the noiseless reference sample on
the GPU (gpu-marked), against the CPU oracle.

SamplerOptions::expected_detectors / expected_observables invert columns in
every row, in detector postselection and in the observable counts
(src/sampler/prepared.rs:909-974, src/sampler/batch/postselect.rs:631-660);
normalize_syndromes takes them from the noiseless reference sample
(src/sampler/prepared.rs:355-405,541-550), which the device computes as one
shot with every noise symbol 0 and the branch stream started at state 1.
Bars: the device reference sample equals the oracle's; rows, counts and
observable counts are bit-exact with the oracle at the per-shot stream
contract on every tier, with explicit bits, with normalization, with
postselection and with pins.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ticit_amd import workloads as w

pytestmark = pytest.mark.gpu

GOLDEN = "X 0\nM 0\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0)\nOBSERVABLE_INCLUDE(1) rec[-1]\n"
MIXED = ("H 0\nT 0\nH 0\nX_ERROR(0.1) 0 1\nM 0 1\nDETECTOR rec[-1] rec[-2]\nH 2\nM 2\nDETECTOR rec[-1]\n"
         "X 3\nDEPOLARIZE1(0.05) 3\nM 3\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-4]\n"
         "H 4\nT 4\nH 4\nX 4\nM 4\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(2) rec[-2] rec[-1]\n")
Q = " ".join(str(q) for q in range(5, 12))
PAIRS = " ".join(f"{q} {q + 1}" for q in range(5, 11))
ENTANGLED = (MIXED + f"H {Q}\nT {Q}\nCX {PAIRS}\nT {Q}\nH {Q}\nDEPOLARIZE1(0.02) {Q}\nM {Q}\n"
             + "".join(f"DETECTOR rec[-{j}] rec[-{j + 1}]\n" for j in range(1, 7))
             + "OBSERVABLE_INCLUDE(1) rec[-7]\n")
CULTIVATION = open(__file__.replace("test_gpu_reference.py", "golden/circuits/msc_d3_inject_cultivate_p1e-3.stim")).read()
CASES = {
    "mixed": MIXED,
    "surface_d3_t": w.surface_code_t(3, 3, 2e-2, 1, 0),
    "cultivation_d3": CULTIVATION,
    "entangled_k7": ENTANGLED,
}


def compile_on(hip, text, tier, *args, **kw):
    c = hip.Circuit(text)
    if tier == "lane" and c.max_active_qubits > 5:
        pytest.skip("the lane tier holds active widths up to 5")
    return c, c.compile(*args, tier=tier, **kw)


@pytest.fixture(autouse=True)
def clear_state():
    yield
    O.set_expected(None, None)
    O.set_pins(None)


def oracle_reference(blob):
    d, o = O.reference_sample(blob)
    return [bool(v) for v in d], [bool(v) for v in o]


def assert_matches(r, o, shots):
    assert (r.shots, r.discarded, r.accepted, r.logical_errors) == tuple(int(v) for v in o["counts"])
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(r, key), o[key]), key
    assert r.observable_ones.tolist() == o["observable_ones"].tolist()


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_device_reference_sample_equals_oracle(hip, name, tier):
    c, s = compile_on(hip, CASES[name], tier)
    ref = s.reference_sample()
    print(name, tier, "max_k", s.max_active_qubits, "reference ones", sum(ref.detectors), flush=True)
    want = oracle_reference(c.blob)
    assert (ref.detectors, ref.observables) == want
    assert (c.reference_sample().detectors, c.reference_sample().observables) == want


# The cultivation circuit's check measurements have probabilities that are
# exactly 0 or 1 in exact arithmetic; the computed value is 0, 1 or within a
# few ulps of them, and the reference's rule draws a uniform only when
# 0 < p < 1. Summation order (a parallel reduction on the group and wide
# tiers, the scalar order on the lane tier and in the oracle) decides that
# draw in about 1 of 2000 shots per measurement, which shifts the shot's
# later draws. There the seeded comparison is against the same tier's raw
# rows; injected uniforms (no stream coupling) give bit-exact rows on every
# tier.
ROUNDING_BOUND = {"cultivation_d3"}


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_normalized_rows_and_counts_match_oracle(hip, name, tier):
    c, s = compile_on(hip, CASES[name], tier, normalize_syndromes=True)
    shots = 2000
    r = s.sample(shots, seed=17, bit_packed=True)
    ref_d, ref_o = oracle_reference(c.blob)
    print(name, tier, "normalized logical errors", r.logical_errors, flush=True)
    # the handle reports the reference before its own inversion
    assert (s.reference_sample().detectors, s.reference_sample().observables) == (ref_d, ref_o)
    raw = c.compile(tier=tier).sample(shots, seed=17, bit_packed=False)
    assert np.array_equal(np.unpackbits(r.detectors, axis=1, bitorder="little")[:, :c.num_detectors],
                          raw.detectors ^ np.array(ref_d, dtype=np.uint8))
    assert np.array_equal(np.unpackbits(r.observables, axis=1, bitorder="little")[:, :len(ref_o)],
                          raw.observables ^ np.array(ref_o, dtype=np.uint8))
    if tier != "lane" and name in ROUNDING_BOUND:
        return
    O.set_expected(ref_d, ref_o)
    o = O.sample_per_shot(c.blob, shots, 17, bit_packed=True)
    assert_matches(r, o, shots)


@pytest.mark.parametrize("tier", ["group", "wide"])
def test_rounding_bound_circuit_injected_rows_match_oracle(hip, tier):
    c = hip.Circuit(CULTIVATION)
    s = c.compile(tier=tier, normalize_syndromes=True)
    shots = 2000
    rng = np.random.default_rng(1)
    words = (shots + 63) // 64
    exo = rng.integers(0, 1 << 62, size=(c.num_symbols, words), dtype=np.uint64) & np.uint64(0x0001000100010001)
    uni = rng.random((s.num_measure_instructions, shots))
    r = s.sample_injected(shots, exo, uni, want_probs=True)
    O.set_expected(*oracle_reference(c.blob))
    o = O.sample_injected(c.blob, shots, exo, uni, None, want_probs=True)
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(r, key), o[key]), key
    assert np.max(np.abs(r.probabilities - o["probs"])) <= 1e-12


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
def test_explicit_bits_match_oracle(hip, tier):
    c = hip.Circuit(MIXED)
    rng = np.random.default_rng(5)
    ed = [bool(v) for v in rng.integers(0, 2, c.num_detectors)]
    eo = [bool(v) for v in rng.integers(0, 2, 3)]
    s = c.compile(tier=tier, expected_detectors=ed, expected_observables=eo, observable=2)
    shots = 1500
    r = s.sample(shots, seed=2, bit_packed=False)
    O.set_expected(ed, eo)
    o = O.sample_per_shot(c.blob, shots, 2, bit_packed=False, observable=2)
    assert_matches(r, o, shots)


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
def test_normalized_postselection_matches_oracle(hip, tier):
    c = hip.Circuit(MIXED)
    mask = [1, 1, 0, 1]
    s = c.compile(mask, tier=tier, normalize_syndromes=True)
    shots = 3000
    r = s.sample(shots, seed=23, bit_packed=True)
    O.set_expected(*oracle_reference(s.blob))
    o = O.sample_per_shot(s.blob, shots, 23, bit_packed=True)
    print(tier, "accepted", r.accepted, "discarded", r.discarded, flush=True)
    assert 0 < r.discarded < shots
    assert_matches(r, o, shots)


def test_reference_fixture_counts(hip):
    """src/sampler/prepared.rs:1164-1194 and ticit_py/tests/test_api.py:56-70."""
    c = hip.Circuit(GOLDEN)
    ref = c.reference_sample()
    assert ref.detectors == [True] and ref.observables == [False, True]
    s = c.compile([1], normalize_syndromes=True, observable=1)
    r = s.sample(4, seed=7)
    assert (r.shots, r.discarded, r.accepted, r.logical_errors) == (4, 0, 4, 0)
    assert r.detectors.tolist() == [[0]] * 4
    assert r.observables.tolist() == [[0, 0]] * 4
    assert r.observable_ones.tolist() == [0, 0]


@pytest.mark.parametrize("value", [False, True])
def test_pinned_reference_normalizes_against_the_pinned_value(hip, value):
    text = "H 0\nM 0\nDETECTOR rec[-1]\nH 1\nT 1\nH 1\nX_ERROR(0.1) 0\nM 1 0\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-1]\n"
    c = hip.Circuit(text)
    pins = [([0], value)]
    s = c.compile(pin_measurements=pins, normalize_syndromes=True)
    ref = s.reference_sample()
    O.set_pins(c.blob, pins)
    assert (ref.detectors, ref.observables) == oracle_reference(c.blob)
    assert ref.detectors[0] == value
    shots = 1000
    r = s.sample(shots, seed=4)
    O.set_expected(*oracle_reference(c.blob))
    o = O.sample_per_shot(c.blob, shots, 4, bit_packed=False)
    assert_matches(r, o, shots)
    assert r.observable_ones[0] < shots // 4  # only the X_ERROR flips the normalized observable


def test_fp32_sampler_normalizes_with_the_fp64_reference(hip):
    """An FP32 sampler takes its reference from an FP64 twin: its normalized
    rows are its own raw rows XOR the oracle's reference."""
    c = hip.Circuit(CASES["entangled_k7"])
    raw = c.compile(precision="fp32")
    s32 = c.compile(normalize_syndromes=True, precision="fp32")
    assert s32.tier == "wide"
    ref_d, ref_o = oracle_reference(c.blob)
    assert (s32.reference_sample().detectors, s32.reference_sample().observables) == (ref_d, ref_o)
    a, b = raw.sample(1000, seed=8), s32.sample(1000, seed=8)
    assert np.array_equal(b.detectors, a.detectors ^ np.array(ref_d, dtype=np.uint8))
    assert np.array_equal(b.observables, a.observables ^ np.array(ref_o, dtype=np.uint8))


def test_device_entry_point_and_batch_apply_the_bits(hip):
    import torch
    texts = [MIXED, CASES["entangled_k7"], CASES["cultivation_d3"]]
    samplers = [hip.Circuit(t).compile(normalize_syndromes=True) for t in texts]
    shots = [700, 500, 300]
    row = lambda n: max(1, (n + 7) // 8)

    def buffers(s, n):
        return (torch.zeros((n, row(s.num_measurements)), dtype=torch.uint8, device="cuda"),
                torch.zeros((n, row(s.num_detectors)), dtype=torch.uint8, device="cuda"),
                torch.zeros((n, row(s.num_observables)), dtype=torch.uint8, device="cuda"),
                torch.zeros(max(1, s.num_observables), dtype=torch.int64, device="cuda"))

    singles = [buffers(s, n) for s, n in zip(samplers, shots)]
    for s, n, b, sd in zip(samplers, shots, singles, (3, 4, 5)):
        s.sample_device(n, sd, b)
    torch.cuda.synchronize()
    for s, n, b, sd in zip(samplers, shots, singles, (3, 4, 5)):
        ref = s.sample(n, seed=sd, bit_packed=True)
        assert np.array_equal(b[1].cpu().numpy(), ref.detectors)
        assert np.array_equal(b[2].cpu().numpy(), ref.observables)
        assert b[3].cpu().numpy()[:s.num_observables].tolist() == ref.observable_ones.tolist()
    outs = [buffers(s, n) for s, n in zip(samplers, shots)]
    bt = hip.Batch(samplers, shots, [3, 4, 5], outs, autotune=False)
    bt.launch(0)
    bt.wait()
    torch.cuda.synchronize()
    for b, o in zip(singles, outs):
        assert all(torch.equal(x, y) for x, y in zip(b, o))


# A postselected detector whose noiseless value is 1, ahead of an entangled
# width-7 block with its own detectors and observables: the reference shot
# must record every later detector (src/sampler/prepared.rs:355-405 runs it
# without postselection), on the tiers that stop a discarded shot early.
POST_FIRST = "X 12\nM 12\nDETECTOR rec[-1]\n" + ENTANGLED


@pytest.mark.parametrize("tier,threads", [("group", 0), ("group", 64), ("group", 256), ("wide", 0)])
def test_reference_shot_runs_past_a_fired_postselected_detector(hip, tier, threads):
    c = hip.Circuit(POST_FIRST)
    mask = [1] + [0] * (c.num_detectors - 1)
    s = c.compile(mask, tier=tier, group_threads=threads, normalize_syndromes=True)
    want = oracle_reference(c.blob)
    assert want[0][0] and sum(want[0]) > 1, want
    ref = s.reference_sample()
    print(tier, threads, "reference ones", sum(ref.detectors), flush=True)
    assert (ref.detectors, ref.observables) == want
    shots = 1500
    r = s.sample(shots, seed=29, bit_packed=True)
    O.set_expected(*want)
    o = O.sample_per_shot(s.blob, shots, 29, bit_packed=True)
    assert_matches(r, o, shots)
    assert r.discarded == 0  # the normalized detector never fires
