"""Detector postselection on the GPU (gpu-marked), against the CPU oracle.

A shot on which a postselected detector fires (DISCARD lines unioned with
the caller's postselection mask, src/circuit/mod.rs:236-243) is discarded
(src/sampler/batch/postselect.rs): it leaves no rows, observable counts or
expectation values, and counts as discarded. Bars: counts, survivor rows
and observable counts bit-exact with the oracle at the per-shot stream
contract, on every tier; injected runs likewise.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ticit_amd import workloads as w

pytestmark = pytest.mark.gpu

NOISY = ("H 0\nT 0\nH 0\nX_ERROR(0.2) 1\nM 0 1\nDETECTOR rec[-1]\nH 2\nT 2\nM 2\n"
         "DETECTOR rec[-1] rec[-2]\nDISCARD rec[-2]\nOBSERVABLE_INCLUDE(0) rec[-3]\n"
         "EXP_VAL Z0\n")
CASES = {
    "noisy_discard": (NOISY, [0, 1]),
    "surface_d3": (w.surface_code_t(3, 3, 2e-2, 1, 0), [1] * 12),
    "c1_first_round": (w.config1_circuit(), [1] * 12),
}


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_survivor_rows_and_counts_match_oracle(hip, name, tier):
    text, mask = CASES[name]
    c = hip.Circuit(text)
    s = c.compile(postselection_mask=mask, tier=tier)
    shots = 3000
    r = s.sample(shots, seed=13, bit_packed=True)
    o = O.sample_per_shot(s.blob, shots, 13, bit_packed=True)
    print(name, tier, "accepted", r.accepted, "discarded", r.discarded, flush=True)
    assert (r.shots, r.accepted, r.discarded) == (shots, int(o["counts"][2]), int(o["counts"][1]))
    assert 0 < r.discarded < shots
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(r, key), o[key]), key
    assert r.observable_ones.tolist() == o["observable_ones"].tolist()
    assert r.logical_errors == int(o["counts"][3])
    if o["exp_vals"].size:
        assert np.max(np.abs(r.exp_vals - o["exp_vals"])) <= 1e-12


def test_injected_postselection_matches_oracle(hip):
    c = hip.Circuit(NOISY)
    s = c.compile(postselection_mask=[0, 1])
    shots = 700
    rng = np.random.default_rng(2)
    words = (shots + 63) // 64
    exo = rng.integers(0, 1 << 62, size=(c.num_symbols, words), dtype=np.uint64)
    uni = rng.random((s.num_measure_instructions, shots))
    r = s.sample_injected(shots, exo, uni)
    o = O.sample_injected(s.blob, shots, exo, uni, None)
    assert (r.accepted, r.discarded) == (int(o["counts"][2]), int(o["counts"][1]))
    assert np.array_equal(r.measurements, o["measurements"])
    assert np.array_equal(r.detectors, o["detectors"])


def test_device_entry_point_refuses_postselected_programs(hip):
    import torch
    s = hip.Circuit(NOISY).compile(postselection_mask=[0, 1])
    b = tuple(torch.zeros((10, 1), dtype=torch.uint8, device="cuda") for _ in range(3)) + (
        torch.zeros(1, dtype=torch.int64, device="cuda"),)
    with pytest.raises(hip.TicitError) as e:
        s.sample_device(10, 1, b)
    assert "postselected" in str(e.value)
