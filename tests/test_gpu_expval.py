"""EXP_VAL probes on the GPU (gpu-marked), against the CPU oracle.

A probe is one of three instructions (src/sampler/batch/runtime.rs:311-356):
a deterministic record (+-1), an unbiased dormant branch (0) or an active
measurement read as sign * (1 - 2 P(true)) with no draw and no collapse
(src/sampler/batch/active.rs:818-850). Bars:
  * rows stay bit-exact with the oracle at the per-shot stream contract:
    probes consume no randomness and leave the state alone;
  * expectation values: lane tier bit-identical (the scalar order), group
    and wide tiers within 1e-12 (fixed FP64 reduction trees);
  * noiseless probes before any measurement equal the exact statevector
    values within 1e-12 on every shot;
  * the device entry point and the batch write the same values.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.statevector import expectation_values
from tests.test_frontend import layered_probe_circuit

pytestmark = pytest.mark.gpu

EXP_TOL = 1e-12


def noisy_probe_circuit(n: int, layers: int, seed: int) -> str:
    """Probes between mid-circuit measurements, resets and noise: the
    probed state depends on every shot's branches."""
    rng = np.random.default_rng(seed)
    base = layered_probe_circuit(n, layers, rng).splitlines()
    out = []
    for line in base[:-1]:
        out.append(line)
        r = rng.random()
        if line.startswith("EXP_VAL") and r < 0.5:
            q = int(rng.integers(n))
            out.append(f"M {q}")
            out.append(f"DETECTOR rec[-1]")
        elif r < 0.7:
            out.append(f"DEPOLARIZE1(0.05) {int(rng.integers(n))}")
    out.append(base[-1])
    out.append("OBSERVABLE_INCLUDE(0) rec[-1]")
    return "\n".join(out) + "\n"


CASES = {
    "layered_n3": layered_probe_circuit(3, 6, np.random.default_rng(1)),
    "layered_n5": layered_probe_circuit(5, 8, np.random.default_rng(2)),
    "noisy_n4": noisy_probe_circuit(4, 8, 3),
    "noisy_n5": noisy_probe_circuit(5, 10, 4),
    "reference_probe": "EXP_VAL Z0 X0\nH 0\nT 0\nEXP_VAL X0 Y0 Z0\nT_DAG 0\nH 0\nM 0\n",
}


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_probe_rows_and_values_match_oracle(hip, name, tier):
    c = hip.Circuit(CASES[name])
    s = c.compile(tier=tier)
    assert s.tier == tier
    shots = 700
    print(name, tier, "max_k", s.max_active_qubits, "probes", s.num_expectation_values, flush=True)
    r = s.sample(shots, seed=21, bit_packed=True)
    o = O.sample_per_shot(c.blob, shots, 21, bit_packed=True)
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(r, key), o[key]), key
    assert r.exp_vals.shape == o["exp_vals"].shape == (shots, s.num_expectation_values)
    err = float(np.max(np.abs(r.exp_vals - o["exp_vals"])))
    print("max expectation error", err)
    if tier == "lane":
        assert np.array_equal(r.exp_vals, o["exp_vals"])
    else:
        assert err <= EXP_TOL


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
def test_noiseless_probes_equal_statevector(hip, tier):
    text = CASES["layered_n5"]
    want = expectation_values(text)
    s = hip.Circuit(text).compile(tier=tier)
    r = s.sample(256, seed=3)
    assert np.max(np.abs(r.exp_vals - want[None, :])) <= EXP_TOL


def test_injected_probe_values_match_oracle(hip):
    c = hip.Circuit(CASES["noisy_n5"])
    s = c.compile()
    shots = 1000
    rng = np.random.default_rng(8)
    words = (shots + 63) // 64
    exo = rng.integers(0, 1 << 62, size=(c.num_symbols, words), dtype=np.uint64) & np.uint64(0x0101010101010101)
    coins = rng.integers(0, 1 << 62, size=(s.num_coin_instructions, words), dtype=np.uint64)
    uni = rng.random((s.num_measure_instructions, shots))
    r = s.sample_injected(shots, exo, uni, coins)
    o = O.sample_injected(c.blob, shots, exo, uni, coins)
    assert np.array_equal(r.measurements, o["measurements"])
    assert np.array_equal(r.exp_vals, o["exp_vals"])  # lane tier: bit-identical


def test_counts_only_returns_no_values(hip):
    s = hip.Circuit(CASES["noisy_n4"]).compile()
    r = s.sample(100, seed=1, keep_records=False)
    assert r.exp_vals.shape == (0, s.num_expectation_values)


def test_device_entry_point_and_batch_write_the_values(hip):
    import torch
    texts = [CASES["noisy_n4"], CASES["layered_n5"], CASES["noisy_n5"]]
    tiers = ["lane", "group", "wide"]
    samplers = [hip.Circuit(t).compile(tier=tier) for t, tier in zip(texts, tiers)]
    shots = [500, 300, 200]
    row = lambda n: max(1, (n + 7) // 8)

    def buffers(s, n):
        return (torch.zeros((n, row(s.num_measurements)), dtype=torch.uint8, device="cuda"),
                torch.zeros((n, row(s.num_detectors)), dtype=torch.uint8, device="cuda"),
                torch.zeros((n, row(s.num_observables)), dtype=torch.uint8, device="cuda"),
                torch.zeros(max(1, s.num_observables), dtype=torch.int64, device="cuda"),
                torch.full((n, s.num_expectation_values), 7.0, dtype=torch.float64, device="cuda"))

    singles = [buffers(s, n) for s, n in zip(samplers, shots)]
    for s, n, b, sd in zip(samplers, shots, singles, (5, 6, 7)):
        s.sample_device(n, sd, b)
    torch.cuda.synchronize()
    for s, n, b, sd in zip(samplers, shots, singles, (5, 6, 7)):
        ref = s.sample(n, seed=sd, bit_packed=True)
        assert np.array_equal(b[4].cpu().numpy(), ref.exp_vals)
    outs = [buffers(s, n) for s, n in zip(samplers, shots)]
    bt = hip.Batch(samplers, shots, [5, 6, 7], outs, autotune=False)
    bt.launch(0)
    bt.wait()
    torch.cuda.synchronize()
    for b, o in zip(singles, outs):
        assert torch.equal(b[0], o[0]) and torch.equal(b[4], o[4])
