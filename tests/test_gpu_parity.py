"""HIP backend parity with the CPU oracle (gpu-marked).

Bars:
  * rows (measurements, detectors, observables) and observable counts are
    bit-exact against the oracle run at the per-shot stream contract
    (reference sampling with one-shot chunks and blocks);
  * with injected randomness, rows are bit-exact and P(branch = 1) agrees to
    1e-12 (lane tier: bit-exact, both run the scalar evaluation order);
  * sampled distributions of small circuits pass a TVD check against exact
    statevector probabilities;
  * at full configuration sizes: counts equal the popcount of the returned
    rows, launches split by first_shot reproduce one launch, and noiseless
    stabilizer circuits keep every detector at zero.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ticit_amd import workloads as w
from tests.statevector import final_distribution, record_values, tvd

pytestmark = pytest.mark.gpu

PROB_TOL = 1e-12


def circuits():
    out = {
        "hth": "H 0\nT 0\nH 0\nM 0\nOBSERVABLE_INCLUDE(0) rec[-1]\n",
        "two_active": "H 0 1\nT 0 1\nH 0 1\nM 0 1\nOBSERVABLE_INCLUDE(0) rec[-1] rec[-2]\n",
        "deterministic": "M !0\nM 1\nDETECTOR rec[-2] rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-2]\n"
                         "OBSERVABLE_INCLUDE(2) rec[-1]\n",
        "surface_d3": w.surface_code_t(3, 3, 1e-2, 1, 0),
        "c1": w.config1_circuit(),
        "c0": w.config0_circuit(),
        "c2_first": w.config2_batch(1)[0][0],
        "noisy_mix": "X_ERROR(0.25) 0\nH 1\nT 1\nH 1\nDEPOLARIZE1(0.1) 1\nY_ERROR(0.5) 2\nM 0 1 2\n"
                     "DETECTOR rec[-1] rec[-2]\nOBSERVABLE_INCLUDE(0) rec[-3]\n",
    }
    return out


CIRCUITS = circuits()


@pytest.mark.parametrize("name", sorted(CIRCUITS))
@pytest.mark.parametrize("packed", [True, False])
def test_seeded_rows_match_oracle(hip, name, packed):
    c = hip.Circuit(CIRCUITS[name])
    s = c.compile()
    shots = 3000
    r = s.sample(shots, seed=11, bit_packed=packed)
    o = O.sample_per_shot(c.blob, shots, 11, bit_packed=packed)
    print(name, "tier", s.tier, "max_k", s.max_active_qubits)
    for key in ("measurements", "detectors", "observables"):
        got, want = getattr(r, key), o[key]
        assert got.shape == want.shape, key
        bad = np.flatnonzero((got != want).any(axis=1)) if got.size else []
        assert len(bad) == 0, f"{name} {key}: {len(bad)} rows differ, first {bad[:5]}"
    assert r.observable_ones.tolist() == o["observable_ones"].tolist()
    assert r.shots == r.accepted == shots and r.discarded == 0


def injected_inputs(s, c, shots, seed):
    rng = np.random.default_rng(seed)
    words = (shots + 63) // 64
    exo = rng.integers(0, 1 << 62, size=(c.num_symbols, words), dtype=np.uint64)
    exo &= np.uint64(0x1111111111111111)  # sparse noise
    coins = rng.integers(0, 1 << 62, size=(s.num_coin_instructions, words), dtype=np.uint64)
    return exo, coins, rng


@pytest.mark.parametrize("name", sorted(CIRCUITS))
def test_injected_rows_and_probabilities_match_oracle(hip, name):
    c = hip.Circuit(CIRCUITS[name])
    s = c.compile()
    shots = 2000
    exo, coins, rng = injected_inputs(s, c, shots, 5)
    # pick uniforms away from the oracle's probabilities so a last-ulp
    # difference in a reduced probability cannot flip a branch
    uni = rng.random((s.num_measure_instructions, shots))
    for _ in range(3):
        o = O.sample_injected(c.blob, shots, exo, uni, coins, want_probs=True)
        close = np.abs(uni - o["probs"]) < 1e-9
        if not close.any():
            break
        uni[close] = np.where(o["probs"][close] < 0.5, o["probs"][close] + 0.25, o["probs"][close] - 0.25)
    r = s.sample_injected(shots, exo, uni, coins, want_probs=True)
    assert np.array_equal(r.measurements, o["measurements"])
    assert np.array_equal(r.detectors, o["detectors"])
    assert np.array_equal(r.observables, o["observables"])
    if o["probs"].size:
        err = np.abs(r.probabilities - o["probs"])
        assert float(err.max()) <= PROB_TOL * max(1.0, float(np.abs(o["probs"]).max()))
        if s.tier == "lane":
            assert np.array_equal(r.probabilities, o["probs"])


def small_circuits():
    out = []
    for i in range(4):
        rng = w.SplitMix64(w.block_seed(0xABCD, 7, i))
        out.append(w.random_clifford_t(3 + i % 3, 6 + i, 2 + i % 4, rng))
    rng = w.SplitMix64(w.block_seed(0xABCD, 8, 0))
    out.append(w.random_clifford_t(8, 10, 7, rng))  # group tier
    return out


@pytest.mark.parametrize("text", small_circuits())
def test_distribution_matches_statevector(hip, text):
    qubits, probs = final_distribution(text)
    s = hip.Circuit(text).compile()
    shots = 200_000
    r = s.sample(shots, seed=3, bit_packed=True)
    vals = record_values(r.measurements, len(qubits))
    counts = np.bincount(vals, minlength=probs.size).astype(float)
    assert tvd(counts, probs) < 0.01
    assert counts[probs < 1e-12].sum() == 0


def test_full_size_c1_properties(hip):
    c = hip.Circuit(w.config1_circuit())
    s = c.compile()
    shots = 1_000_000
    r = s.sample(shots, seed=1, bit_packed=True)
    assert r.measurements.shape == (shots, 19) and r.detectors.shape == (shots, 15)
    ones = int(np.unpackbits(r.observables, axis=1, bitorder="little")[:, 0].sum())
    assert ones == int(r.observable_ones[0]) == r.logical_errors
    # a prefix of a big launch equals a small launch with the same seed
    small = s.sample(5000, seed=1, bit_packed=True)
    assert np.array_equal(small.measurements, r.measurements[:5000])
    # detector rate of a d=5, p=1e-3 memory with T gates stays in a sane band
    rate = np.unpackbits(r.detectors, axis=1, bitorder="little")[:, :120].mean()
    assert 0.005 < rate < 0.06


def test_noiseless_surface_code_full_size(hip):
    s = hip.Circuit(w.surface_code_t(5, 5, 0.0, 0, 0)).compile()
    r = s.sample(1_000_000, seed=2, bit_packed=True)
    assert int(np.unpackbits(r.detectors, axis=1, bitorder="little").sum()) == 0
    assert int(r.observable_ones[0]) == 0


def test_device_entry_point_split_launches(hip):
    import torch
    c = hip.Circuit(w.config1_circuit())
    s = c.compile()
    shots, half = 20000, 7000
    meas = torch.zeros((shots, 19), dtype=torch.uint8, device="cuda")
    det = torch.zeros((shots, 15), dtype=torch.uint8, device="cuda")
    obs = torch.zeros((shots, 1), dtype=torch.uint8, device="cuda")
    ones = torch.zeros(1, dtype=torch.int64, device="cuda")
    lib = hip._capi.load()
    for first, n in ((0, half), (half, shots - half)):
        hip._capi.check(lib.ticit_hip_sample_device(
            s.handle, n, 4, first, 1, 1, meas[first:].data_ptr(), det[first:].data_ptr(),
            obs[first:].data_ptr(), ones.data_ptr(), None, None))
    torch.cuda.synchronize()
    ref = s.sample(shots, seed=4, bit_packed=True)
    assert np.array_equal(meas.cpu().numpy(), ref.measurements)
    assert np.array_equal(det.cpu().numpy(), ref.detectors)
    assert int(ones.item()) == int(ref.observable_ones[0])


@pytest.mark.parametrize("tps", [8, 16, 32, 64, 128, 256, 512])
def test_group_threads_per_shot_give_identical_rows(hip, tps):
    """Packing 8, 4 or 2 shots per wave, or one shot per workgroup of 64 to
    512 threads, changes the launch shape, not the rows: they equal the
    oracle's, and P(branch = 1) agrees to 1e-12 (the tier's pairwise tree
    against the oracle's scalar order)."""
    c = hip.Circuit(w.config2_batch()[94][0])  # component layout, local k 6
    s = c.compile(tier="group", group_threads=tps)
    shots = 333  # not a multiple of any shots-per-workgroup
    r = s.sample(shots, seed=5, bit_packed=True)
    o = O.sample_per_shot(c.blob, shots, 5)
    assert np.array_equal(r.measurements, o["measurements"])
    assert r.observable_ones.tolist() == o["observable_ones"].tolist()
    rng = np.random.default_rng(tps)
    uni = rng.random((s.num_measure_instructions, shots))
    coins = rng.integers(0, 1 << 62, size=(max(1, s.num_coin_instructions), (shots + 63) // 64), dtype=np.uint64)
    ri = s.sample_injected(shots, uniforms=uni, coins=coins, want_probs=True)
    oi = O.sample_injected(c.blob, shots, None, uni, coins, want_probs=True)
    assert np.array_equal(ri.measurements, oi["measurements"])
    assert np.max(np.abs(ri.probabilities - oi["probs"])) <= PROB_TOL


def test_group_tier_sums_do_not_depend_on_threads_per_shot(hip):
    """The group tier sums a probability pairwise in index order at every
    threads-per-shot, so P(branch = 1) is bit-identical across launch shapes
    and so are seeded rows, also where a deterministic measurement's sum
    lands on exactly 1 in one order and one ulp below in another (the draw
    of a uniform depends on it). Planner-corpus program 100321 (components,
    local k 8) has such measurements: with per-thread sums in strided order,
    16 threads per shot gave 2 of 3000 seeded rows unlike 8, 32 or 64."""
    c = hip.Circuit(w.planner_corpus_circuit(100321))
    shots = 3000
    rows, probs = {}, {}
    rng = np.random.default_rng(3)
    uni = None
    for tps in (8, 16, 32, 64, 128, 256):
        s = c.compile(tier="group", group_threads=tps, components="components")
        if uni is None:
            uni = rng.random((s.num_measure_instructions, shots))
        rows[tps] = s.sample(shots, seed=6, bit_packed=True).measurements
        probs[tps] = s.sample_injected(shots, uniforms=uni, want_probs=True).probabilities
    for tps in rows:
        assert np.array_equal(rows[tps], rows[8]), tps
        assert np.array_equal(probs[tps], probs[8]), tps


def test_group_tier_holds_a_width_13_state(hip):
    """wide_compact (max_k 13, a 128 KiB FP64 state) runs in LDS with 512
    threads per shot when asked for: rows bit-exact, P(branch = 1) within
    1e-12."""
    text = open(__file__.replace("test_gpu_parity.py", "golden/circuits/wide_compact.stim")).read()
    c = hip.Circuit(text)
    s = c.compile(tier="group", group_threads=512)
    assert s.tier == "group" and s.max_active_qubits == 13
    shots = 500
    r = s.sample(shots, seed=9, bit_packed=True)
    o = O.sample_per_shot(c.blob, shots, 9)
    assert np.array_equal(r.measurements, o["measurements"])
    rng = np.random.default_rng(4)
    uni = rng.random((s.num_measure_instructions, shots))
    ri = s.sample_injected(shots, uniforms=uni, want_probs=True)
    oi = O.sample_injected(c.blob, shots, None, uni, None, want_probs=True)
    assert np.array_equal(ri.measurements, oi["measurements"])
    assert np.max(np.abs(ri.probabilities - oi["probs"])) <= PROB_TOL


def chain_circuit(n: int) -> str:
    """H T on every qubit, a CX chain, H T again, measure all: a dense active
    state of width n."""
    lines = [f"H {q}\nT {q}" for q in range(n)] + [f"CX {q} {q + 1}" for q in range(n - 1)]
    lines += [f"H {q}\nT {q}" for q in range(n)] + ["M " + " ".join(map(str, range(n)))]
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("n,tier", [(10, "group"), (11, "wide")])
def test_automatic_tier_keeps_the_group_tier_up_to_width_10(hip, n, tier):
    """The automatic choice runs dense widths up to 10 in the group tier and
    wider ones in the wide tier. An explicit group request still runs width
    11, at 256 threads per shot; every tier's rows equal the oracle's."""
    c = hip.Circuit(chain_circuit(n))
    s = c.compile(components="dense")
    assert s.max_local_k == n and s.tier == tier
    shots = 300
    o = O.sample_per_shot(c.blob, shots, 11)
    for smp in (s, c.compile(components="dense", tier="group"), c.compile(components="dense", tier="wide")):
        r = smp.sample(shots, seed=11, bit_packed=True)
        assert np.array_equal(r.measurements, o["measurements"]), smp.tier
