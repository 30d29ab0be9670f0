"""Detector postselection on the GPU (gpu-marked), against the CPU oracle.

A shot on which a postselected detector fires (DISCARD lines unioned with
the caller's postselection mask, src/circuit/mod.rs:236-243) is discarded
(src/sampler/batch/postselect.rs): it leaves no rows, observable counts or
expectation values, and counts as discarded. Bars: counts, survivor rows
and observable counts bit-exact with the oracle at the per-shot stream
contract, on every tier; injected runs likewise. Survivors are compacted on
the device: ticit_hip_sample, ticit_hip_sample_device_postselected (rows at
the front of the caller's buffers, the count in a device word) and the batch
(compaction captured in its graph) give the same rows.
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


def test_plain_device_entry_point_refuses_postselected_programs(hip):
    import torch
    s = hip.Circuit(NOISY).compile(postselection_mask=[0, 1])
    b = tuple(torch.zeros((10, 1), dtype=torch.uint8, device="cuda") for _ in range(3)) + (
        torch.zeros(1, dtype=torch.int64, device="cuda"),)
    with pytest.raises(hip.TicitError) as e:
        s.sample_device(10, 1, b)
    assert "ticit_hip_sample_device_postselected" in str(e.value)


def device_buffers(s, n, fill=0xAB):
    import torch
    row = lambda c: max(1, (c + 7) // 8)
    return (torch.full((n, row(s.num_measurements)), fill, dtype=torch.uint8, device="cuda"),
            torch.full((n, row(s.num_detectors)), fill, dtype=torch.uint8, device="cuda"),
            torch.full((n, row(s.num_observables)), fill, dtype=torch.uint8, device="cuda"),
            torch.zeros(max(1, s.num_observables), dtype=torch.int64, device="cuda"),
            torch.full((n, s.num_expectation_values), 7.0, dtype=torch.float64, device="cuda"),
            torch.full((1,), -1, dtype=torch.int64, device="cuda"))


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_device_survivors_match_host_survivors(hip, name, tier):
    """ticit_hip_sample_device_postselected: survivors' rows at the front of
    the caller's buffers, their number in the accepted word; two calls in a
    row (the second smaller) reuse the program's staging."""
    import torch
    text, mask = CASES[name]
    s = hip.Circuit(text).compile(postselection_mask=mask, tier=tier)
    for shots, seed in ((3000, 13), (1100, 14)):
        b = device_buffers(s, shots)
        s.sample_device(shots, seed, b[:5], accepted=b[5])
        torch.cuda.synchronize()
        ref = s.sample(shots, seed=seed, bit_packed=True)
        acc = int(b[5].item())
        print(name, tier, shots, "accepted", acc, flush=True)
        assert acc == ref.accepted and 0 < acc < shots
        assert np.array_equal(b[0][:acc].cpu().numpy(), ref.measurements)
        assert np.array_equal(b[1][:acc].cpu().numpy(), ref.detectors)
        assert np.array_equal(b[2][:acc].cpu().numpy(), ref.observables)
        assert np.array_equal(b[4][:acc].cpu().numpy(), ref.exp_vals)
        assert b[3].cpu().numpy()[:s.num_observables].tolist() == ref.observable_ones.tolist()
        # rows past the survivors are untouched
        assert bool((b[0][acc:] == 0xAB).all())


@pytest.mark.parametrize("group_threads", [8, 64, 256])
def test_group_tier_shapes_agree_under_postselection(hip, group_threads):
    """One shot per workgroup stops a shot at its first fired postselected
    detector; packed sub-wave shots run on. Both give the oracle's
    survivors."""
    text, mask = CASES["surface_d3"]
    s = hip.Circuit(text).compile(postselection_mask=mask, tier="group", group_threads=group_threads)
    shots = 2500
    r = s.sample(shots, seed=31, bit_packed=True)
    o = O.sample_per_shot(s.blob, shots, 31, bit_packed=True)
    assert (r.accepted, r.discarded) == (int(o["counts"][2]), int(o["counts"][1]))
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(r, key), o[key]), key
    assert r.observable_ones.tolist() == o["observable_ones"].tolist()


def test_counts_only_and_edge_survivor_counts(hip):
    import torch
    always = hip.Circuit("X 0\nM 0\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-1]\n").compile([1])
    r = always.sample(5000, seed=1, keep_records=False)
    assert (r.accepted, r.discarded, r.logical_errors) == (0, 5000, 0)
    b = device_buffers(always, 5000)
    always.sample_device(5000, 1, b[:4], accepted=b[5])
    torch.cuda.synchronize()
    assert int(b[5].item()) == 0 and bool((b[0] == 0xAB).all())
    never = hip.Circuit("M 0\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-1]\n").compile([1])
    b = device_buffers(never, 4097)
    never.sample_device(4097, 1, b[:4], accepted=b[5])
    torch.cuda.synchronize()
    assert int(b[5].item()) == 4097 and bool((b[0] == 0).all())
    plain = hip.Circuit(CASES["surface_d3"][0]).compile()  # no DISCARD, no mask
    assert not plain.has_postselection
    b = device_buffers(plain, 300)
    plain.sample_device(300, 2, b[:5], accepted=b[5])
    torch.cuda.synchronize()
    assert int(b[5].item()) == 300
    ref = plain.sample(300, seed=2, bit_packed=True)
    assert np.array_equal(b[0].cpu().numpy(), ref.measurements)


def test_batch_compacts_postselected_items(hip):
    import torch
    samplers = [hip.Circuit(CASES[n][0]).compile(postselection_mask=CASES[n][1], tier=t)
                for n, t in (("noisy_discard", "lane"), ("surface_d3", "group"), ("c1_first_round", "wide"))]
    samplers.append(hip.Circuit(CASES["surface_d3"][0]).compile())  # no postselection in the same batch
    shots = [3000, 2000, 700, 500]
    seeds = [3, 4, 5, 6]
    outs = [device_buffers(s, n) for s, n in zip(samplers, shots)]
    bt = hip.Batch(samplers, shots, seeds, outs, autotune=False)
    bt.launch(0)
    bt.wait()
    torch.cuda.synchronize()
    for s, n, sd, b in zip(samplers, shots, seeds, outs):
        ref = s.sample(n, seed=sd, bit_packed=True)
        acc = int(b[5].item()) if s.has_postselection else n
        assert acc == ref.accepted
        assert np.array_equal(b[0][:acc].cpu().numpy(), ref.measurements)
        assert np.array_equal(b[1][:acc].cpu().numpy(), ref.detectors)
        assert np.array_equal(b[4][:acc].cpu().numpy(), ref.exp_vals)
        assert b[3].cpu().numpy()[:s.num_observables].tolist() == ref.observable_ones.tolist()
    res = hip.sample_batch(samplers, shots, seeds, bit_packed=True)
    for s, n, sd, r in zip(samplers, shots, seeds, res):
        ref = s.sample(n, seed=sd, bit_packed=True)
        assert (r.accepted, r.discarded) == (ref.accepted, ref.discarded)
        assert np.array_equal(r.measurements, ref.measurements)
        assert np.array_equal(r.exp_vals, ref.exp_vals)


def test_wide_tier_skips_discarded_shots(hip):
    """A postselected detector that fires in about 90 % of shots sits before
    a width-12 section (14 qubits, 8 layers): the wide tier stops those shots at the detector
    (src/sampler/batch/postselect.rs removes them), and the survivors' rows
    and counts still equal the oracle's."""
    rng = np.random.default_rng(12)
    lines = ["X_ERROR(0.9) 20", "M 20", "DETECTOR rec[-1]"]
    for layer in range(8):
        for q in range(14):
            lines.append(f"{rng.choice(['H', 'S'])} {q}")
            if rng.random() < 0.5:
                lines.append(f"T {q}")
        for q in range(layer % 2, 13, 2):
            lines.append(f"CX {q} {q + 1}")
    lines.append("M " + " ".join(str(q) for q in range(14)))
    lines.append("OBSERVABLE_INCLUDE(0) rec[-1] rec[-2]")
    text = "\n".join(lines) + "\n"
    c = hip.Circuit(text)
    s = c.compile([1], tier="wide")
    assert s.tier == "wide" and s.max_active_qubits >= 10
    shots = 400
    r = s.sample(shots, seed=3, bit_packed=True)
    o = O.sample_per_shot(s.blob, shots, 3, bit_packed=True)
    print("accepted", r.accepted, "of", shots, "max_k", s.max_active_qubits, flush=True)
    assert (r.accepted, r.discarded) == (int(o["counts"][2]), int(o["counts"][1]))
    assert 0 < r.accepted < shots // 4
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(r, key), o[key]), key
    assert r.observable_ones.tolist() == o["observable_ones"].tolist()


def test_batch_replays_match_eager_runs_between_them(hip):
    """Four replays of one captured batch (lane, group and wide items, three
    of them postselected with their compaction inside the graph), with eager
    sampling and synchronization between the replays: every replay's
    survivors, rows and counts equal an eager run at the same seeds."""
    import torch
    samplers = [hip.Circuit(CASES[n][0]).compile(postselection_mask=CASES[n][1], tier=t)
                for n, t in (("noisy_discard", "lane"), ("surface_d3", "group"), ("c1_first_round", "wide"))]
    samplers.append(hip.Circuit(CASES["surface_d3"][0]).compile())
    shots = [3000, 2000, 700, 500]
    seeds = [3, 4, 5, 6]
    outs = [device_buffers(s, n) for s, n in zip(samplers, shots)]
    bt = hip.Batch(samplers, shots, seeds, outs, autotune=True, autotune_shots=256)
    for off in (0, 5, 5, 12):
        for b in outs:
            b[3].zero_()
        bt.launch(off)
        bt.wait()
        torch.cuda.synchronize()
        print("replay", off, flush=True)
        for s, n, sd, b in zip(samplers, shots, seeds, outs):
            ref = s.sample(n, seed=sd + off, bit_packed=True)
            acc = int(b[5].item()) if s.has_postselection else n
            assert acc == ref.accepted, (off, s.tier)
            assert np.array_equal(b[0][:acc].cpu().numpy(), ref.measurements), (off, s.tier)
            assert np.array_equal(b[1][:acc].cpu().numpy(), ref.detectors), (off, s.tier)
            assert b[3].cpu().numpy()[:s.num_observables].tolist() == ref.observable_ones.tolist(), (off, s.tier)
    bt.close()
