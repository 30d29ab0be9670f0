"""Lane tier with its tables in device memory (gpu-marked).

A lane-tier wave keeps its bit tables [item][lane] in LDS when they fit. A
program whose tables do not fit (the reference's CCZ fixtures: 369 KiB of
bit tables per 64 shots for d05) keeps them in a per-wave device memory
region with the same layout, and above width 5 its amplitudes too. The
kernel is the same interpreter with the same scalar evaluation order, so:

  * seeded rows are bit-exact against the oracle at the per-shot stream
    contract, and injected probabilities are bit-identical to the oracle's;
  * every placement of the same program gives the same rows, including a
    launch with more waves than resident regions (the kernel strides its
    wave blocks over the regions);
  * postselection with normalized syndromes, and a batch item in device
    memory, reproduce the LDS placement and a plain call.
"""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ticit_amd import workloads as w
from tests.test_gpu_parity import CIRCUITS, injected_inputs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "circuits")
DEV = torch.device("cuda:0")


def wide_lane_circuits():
    """Dense programs of widths 6, 7 and 8 (seeded C0-style generator, n = 12,
    16 layers), one with DEPOLARIZE1 and mid-circuit measurements."""
    out = {}
    for k, i, t in ((6, 11, 9), (7, 3, 11), (8, 31, 9)):
        rng = w.SplitMix64(w.block_seed(0xABCD, 10, i))
        out[f"width{k}"] = w.random_clifford_t(12, 16, t, rng)
    rng = w.SplitMix64(w.block_seed(0xABCD, 10, 3))
    out["width6_noisy"] = w.random_clifford_t(12, 16, 11, rng, depolarize1=0.02, mid_measure={5: [0, 3], 11: [7]})
    return out


WIDE = wide_lane_circuits()
NARROW = {k: CIRCUITS[k] for k in ("hth", "two_active", "deterministic", "surface_d3", "c1", "noisy_mix")}


def check_rows(r, o, label):
    for key in ("measurements", "detectors", "observables"):
        got, want = getattr(r, key), o[key]
        assert got.shape == want.shape, (label, key)
        bad = np.flatnonzero((got != want).any(axis=1)) if got.size else []
        assert len(bad) == 0, f"{label} {key}: {len(bad)} rows differ, first {bad[:5]}"
    assert r.observable_ones.tolist() == o["observable_ones"].tolist(), label


@pytest.mark.parametrize("name", sorted(NARROW))
def test_bit_tables_in_device_memory_match_oracle(hip, name):
    c = hip.Circuit(NARROW[name])
    s = c.compile(tier="lane", lane_store="bits_global")
    assert s.tier == "lane" and s.lane_store == "bits_global" and s.lane_region_bytes > 0
    shots = 3000
    r = s.sample(shots, seed=11, bit_packed=True)
    check_rows(r, O.sample_per_shot(c.blob, shots, 11, bit_packed=True), name)


@pytest.mark.parametrize("name", sorted(WIDE))
@pytest.mark.parametrize("packed", [True, False])
def test_all_tables_in_device_memory_match_oracle(hip, name, packed):
    c = hip.Circuit(WIDE[name])
    s = c.compile(tier="lane")
    print(name, "max_k", s.max_active_qubits, "store", s.lane_store, "region", s.lane_region_bytes, flush=True)
    assert s.tier == "lane" and s.lane_store == "all_global" and 6 <= s.max_active_qubits <= 8
    shots = 2000
    r = s.sample(shots, seed=23, bit_packed=packed)
    check_rows(r, O.sample_per_shot(c.blob, shots, 23, bit_packed=packed), name)


@pytest.mark.parametrize("name", sorted(WIDE) + ["surface_d3", "noisy_mix"])
def test_injected_probabilities_are_the_oracles(hip, name):
    text = WIDE.get(name) or NARROW[name]
    c = hip.Circuit(text)
    s = c.compile(tier="lane", lane_store="all_global" if name in WIDE else "bits_global")
    shots = 1000
    exo, coins, rng = injected_inputs(s, c, shots, 9)
    uni = rng.random((s.num_measure_instructions, shots))
    o = O.sample_injected(c.blob, shots, exo, uni, coins, want_probs=True)
    r = s.sample_injected(shots, exo, uni, coins, want_probs=True)
    assert np.array_equal(r.measurements, o["measurements"])
    assert np.array_equal(r.detectors, o["detectors"])
    assert np.array_equal(r.observables, o["observables"])
    if o["probs"].size:
        assert np.array_equal(r.probabilities, o["probs"])  # the same scalar order


def device_buffers(s, shots):
    row = lambda n: max(1, (n + 7) // 8)
    return (torch.zeros((shots, row(s.num_measurements)), dtype=torch.uint8, device=DEV),
            torch.zeros((shots, row(s.num_detectors)), dtype=torch.uint8, device=DEV),
            torch.zeros((shots, row(s.num_observables)), dtype=torch.uint8, device=DEV),
            torch.zeros(max(1, s.num_observables), dtype=torch.int64, device=DEV))


def test_more_waves_than_regions_equal_the_lds_placement(hip):
    # 600k shots = 9375 waves: more than the resident regions, so every
    # region runs several wave blocks in turn
    c = hip.Circuit(w.config1_circuit())
    lds = c.compile()
    dev = c.compile(tier="lane", lane_store="bits_global")
    assert lds.lane_store == "lds" and dev.lane_store == "bits_global"
    shots = 600_000
    a, b = device_buffers(lds, shots), device_buffers(dev, shots)
    lds.sample_device(shots, 5, a)
    dev.sample_device(shots, 5, b)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    # and a second, smaller call on the grown regions
    small = device_buffers(dev, 777)
    dev.sample_device(777, 5, small)
    torch.cuda.synchronize()
    for x, y in zip(small[:3], b[:3]):
        assert torch.equal(x, y[:777])


def test_postselected_normalized_rows_equal_the_lds_placement(hip):
    with open(os.path.join(GOLDEN, "msc_d3_inject_cultivate_p1e-3.stim")) as f:
        c = hip.Circuit(f.read())
    mask = [1] * c.num_detectors
    lds = c.compile(mask, normalize_syndromes=True)
    dev = c.compile(mask, normalize_syndromes=True, tier="lane", lane_store="bits_global")
    assert lds.tier == "lane" and lds.lane_store == "lds" and dev.lane_store == "bits_global"
    assert dev.reference_sample().detectors == lds.reference_sample().detectors
    ra = lds.sample(20000, seed=3, bit_packed=True)
    rb = dev.sample(20000, seed=3, bit_packed=True)
    assert (ra.accepted, ra.discarded) == (rb.accepted, rb.discarded) and 0 < rb.accepted < 20000
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(ra, key), getattr(rb, key)), key


def test_ccz_d05_runs_in_the_lane_tier_in_device_memory(hip):
    with open(os.path.join(GOLDEN, "ccz_d05_p1e-3.clifft")) as f:
        c = hip.Circuit(f.read())
    s = c.compile(normalize_syndromes=True)
    assert (s.tier, s.lane_store, s.max_active_qubits) == ("lane", "all_global", 8)
    # rows and expectation values against the oracle are checked in
    # test_gpu_channels.py; here a call's prefix equals a shorter call
    r = s.sample(4096, seed=41, bit_packed=True)
    assert r.detectors.shape == (4096, (s.num_detectors + 7) // 8)
    short = s.sample(256, seed=41, bit_packed=True)
    assert np.array_equal(short.detectors, r.detectors[:256])
    assert np.array_equal(short.exp_vals, r.exp_vals[:256])


def test_batch_item_in_device_memory_equals_a_plain_call(hip):
    with open(os.path.join(GOLDEN, "ccz_d05_p1e-3.clifft")) as f:
        ccz = hip.Circuit(f.read()).compile(normalize_syndromes=True)
    c1 = hip.Circuit(w.config1_circuit()).compile()
    wk = hip.Circuit(WIDE["width8"]).compile()
    samplers, shots, seeds = [ccz, c1, wk, ccz], [3000, 5000, 2500, 1000], [7, 8, 9, 10]
    outs = [device_buffers(s, n) for s, n in zip(samplers, shots)]
    b = hip.Batch(samplers, shots, seeds, outs, autotune=False)
    b.launch(0)
    b.wait()
    for s, n, sd, out in zip(samplers, shots, seeds, outs):
        ref = device_buffers(s, n)
        s.sample_device(n, sd, ref)
        torch.cuda.synchronize()
        for x, y in zip(out, ref):
            assert torch.equal(x, y), (s.tier, s.lane_store)


def test_placement_requests_are_checked(hip):
    c1 = hip.Circuit(w.config1_circuit())
    with pytest.raises(hip.TicitError):
        c1.compile(tier="lane", lane_store="all_global")  # width 2: amplitudes stay in LDS
    wide = hip.Circuit(WIDE["width7"])
    with pytest.raises(hip.TicitError):
        wide.compile(tier="lane", lane_store="lds")
    with pytest.raises(hip.TicitError):
        wide.compile(tier="lane", lane_store="bits_global")
