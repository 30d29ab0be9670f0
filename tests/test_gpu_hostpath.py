"""Host-result calls that return rows in chunks (gpu-marked).

ticit_hip_sample (Sampler::sample_with_seed, src/sampler/prepared.rs:663-670)
returns the rows of lane- and group-tier programs in chunks of at least
65536 shots: each chunk is sampled and copied to pinned staging while the
host moves the previous one into the result. These tests take enough shots
for several chunks with a ragged last one and check every row, expectation
value and observable count against one launch of ticit_hip_sample_device at
the same seed, whose rows the parity tests pin to the oracle. Repeated calls
on one program reuse its device buffers and staging, so a second, smaller
call must not see the first one's rows.
"""
import numpy as np
import pytest
import torch

from tests.test_gpu_expval import noisy_probe_circuit

pytestmark = pytest.mark.gpu

CHUNK = 65536


def device_rows(s, shots: int, seed: int, packed: bool):
    dev = torch.device("cuda:0")
    cols = (lambda n: (n + 7) // 8) if packed else (lambda n: n)
    meas = torch.empty((shots, cols(s.num_measurements)), dtype=torch.uint8, device=dev)
    det = torch.empty((shots, cols(s.num_detectors)), dtype=torch.uint8, device=dev)
    obs = torch.empty((shots, cols(s.num_observables)), dtype=torch.uint8, device=dev)
    ones = torch.zeros(max(1, s.num_observables), dtype=torch.int64, device=dev)
    outs = [meas, det, obs, ones]
    if s.num_expectation_values:
        outs.append(torch.empty((shots, s.num_expectation_values), dtype=torch.float64, device=dev))
    s.sample_device(shots, seed, outs, bit_packed=packed)
    torch.cuda.synchronize()
    host = [t.cpu().numpy() for t in outs]
    host[3] = host[3][: s.num_observables].astype(np.uint64)
    return host


def check_same(s, shots: int, seed: int, packed: bool):
    r = s.sample(shots, seed, bit_packed=packed)
    want = device_rows(s, shots, seed, packed)
    assert r.shots == shots and r.accepted == shots
    assert np.array_equal(r.measurements, want[0])
    assert np.array_equal(r.detectors, want[1])
    assert np.array_equal(r.observables, want[2])
    assert np.array_equal(r.observable_ones, want[3])
    if s.num_expectation_values:
        assert np.array_equal(r.exp_vals, want[4])
    return r


@pytest.mark.parametrize("packed", [True, False])
def test_surface_code_rows_in_chunks_match_one_launch(hip, packed):
    from ticit_amd import workloads as w
    s = hip.Circuit(w.config1_circuit()).compile()
    assert s.tier == "lane"
    shots = 3 * CHUNK + 1000
    check_same(s, shots, 11, packed)
    # a smaller call on the same program, after the larger one
    check_same(s, CHUNK + 77, 12, packed)


@pytest.mark.parametrize("tier", ["lane", "group"])
def test_probe_rows_and_values_in_chunks_match_one_launch(hip, tier):
    s = hip.Circuit(noisy_probe_circuit(4, 8, 3)).compile(tier=tier)
    assert s.num_expectation_values > 0
    r = check_same(s, 2 * CHUNK + 4097, 21, False)
    # the observable is a noisy measurement: rows differ between shots
    assert 0 < int(r.observable_ones[0]) < r.shots


def test_counts_only_and_repeated_calls_agree(hip):
    from ticit_amd import workloads as w
    s = hip.Circuit(w.config1_circuit()).compile()
    shots = 2 * CHUNK + 5
    full = s.sample(shots, 31, bit_packed=True)
    counts = s.sample(shots, 31, bit_packed=True, keep_records=False)
    again = s.sample(shots, 31, bit_packed=True)
    assert np.array_equal(full.observable_ones, counts.observable_ones)
    assert counts.measurements.shape[0] == 0
    assert np.array_equal(full.measurements, again.measurements)
    assert np.array_equal(full.detectors, again.detectors)
