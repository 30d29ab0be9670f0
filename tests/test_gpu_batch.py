"""Batched sampling (ticit_hip_batch_*, ticit_hip_sample_batch; gpu-marked).

A batch is many single-program sampling calls run as one captured graph, so
each item's rows must equal ticit_hip_sample_device for the same program,
shots and seed, whatever launch shape the autotuner picked (threads per shot
change only the summation tree of a shot's probability sums, which leaves
the rows unchanged on these inputs).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from ticit_amd import workloads as W

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def buffers(s, shots, packed=True):
    row = (lambda n: (n + 7) // 8) if packed else (lambda n: n)
    mk = lambda n: torch.zeros((shots, max(1, row(n))), dtype=torch.uint8, device=DEV)
    return (mk(s.num_measurements), mk(s.num_detectors), mk(s.num_observables),
            torch.zeros(max(1, s.num_observables), dtype=torch.int64, device=DEV))


def mixed_set(hip):
    """C2 circuits spanning every tier and both layouts, shots trimmed."""
    batch = W.config2_batch()
    picks = list(range(0, 48)) + [94, 214, 235, 212, 250]
    samplers, shots, seeds = [], [], []
    for j, i in enumerate(picks):
        s = hip.Circuit(batch[i][0]).compile()
        samplers.append(s)
        shots.append(400 if s.tier == "wide" else 1000 + 37 * j)
        seeds.append(1000 + i)
    return samplers, shots, seeds


def test_batch_rows_equal_single_program_runs(hip):
    samplers, shots, seeds = mixed_set(hip)
    tiers = {s.tier for s in samplers}
    assert tiers == {"lane", "group", "wide"}, tiers
    assert any(s.components for s in samplers) and not all(s.components for s in samplers)
    outs = [buffers(s, n) for s, n in zip(samplers, shots)]
    b = hip.Batch(samplers, shots, seeds, outs, autotune=True, autotune_shots=256)
    b.launch(0)
    b.wait()
    info = b.info()
    print(info, flush=True)
    assert info["items"] == len(samplers) and info["wide_items"] > 0 and info["buckets"]
    for s, n, sd, out in zip(samplers, shots, seeds, outs):
        ref = buffers(s, n)
        s.sample_device(n, sd, ref)
        torch.cuda.synchronize()
        for a, r in zip(out, ref):
            assert torch.equal(a, r), (s.tier, s.components)


def test_batch_replays_take_the_seed_offset(hip):
    samplers, shots, seeds = mixed_set(hip)
    samplers, shots, seeds = samplers[40:], shots[40:], seeds[40:]
    outs = [buffers(s, n) for s, n in zip(samplers, shots)]
    b = hip.Batch(samplers, shots, seeds, outs, autotune=False)
    b.launch(0)
    b.launch(7)
    b.wait()
    for s, n, sd, out in zip(samplers, shots, seeds, outs):
        ref = buffers(s, n)
        s.sample_device(n, sd, ref)        # the first replay's counts
        s.sample_device(n, sd + 7, ref)    # rows and counts of the second
        torch.cuda.synchronize()
        for a, r in zip(out, ref):
            assert torch.equal(a, r)


def test_sample_batch_host_results_match_the_oracle(hip):
    batch = W.config2_batch()
    samplers = [hip.Circuit(batch[i][0]).compile() for i in (3, 94, 235)]
    shots = [700, 500, 150]
    seeds = [11, 12, 13]
    res = hip.sample_batch(samplers, shots, seeds, bit_packed=True)
    for s, n, sd, r in zip(samplers, shots, seeds, res):
        o = O.sample_per_shot(s.blob, n, sd)
        assert np.array_equal(r.measurements, o["measurements"])
        assert r.shots == n


def test_small_width_11_program_joins_a_group_bucket(hip):
    """A dense width-11 program runs wide on its own; in a batch it joins a
    group bucket when it has at most 4e5 shots and stays wide above. Both
    ways its rows equal the single-program run's and the oracle's."""
    n = 11
    lines = [f"H {q}\nT {q}" for q in range(n)] + [f"CX {q} {q + 1}" for q in range(n - 1)]
    lines += [f"H {q}\nT {q}" for q in range(n)] + ["M " + " ".join(map(str, range(n)))]
    c = hip.Circuit("\n".join(lines) + "\n")
    s = c.compile(components="dense")
    assert s.tier == "wide" and s.max_local_k == n
    for shots, wide in ((3000, 0), (400001, 1)):
        out = buffers(s, shots)
        b = hip.Batch([s], [shots], [21], [out], autotune=True, autotune_shots=256)
        b.launch(0)
        b.wait()
        info = b.info()
        print(shots, info, flush=True)
        assert info["wide_items"] == wide and len(info["buckets"]) == 1 - wide
        ref = buffers(s, shots)
        s.sample_device(shots, 21, ref)
        torch.cuda.synchronize()
        for a, r in zip(out, ref):
            assert torch.equal(a, r)
        b.close()
    o = O.sample_per_shot(c.blob, 3000, 21)
    r = s.sample(3000, seed=21, bit_packed=True)
    assert np.array_equal(r.measurements, o["measurements"])


def dense_chain(hip, n):
    lines = [f"H {q}\nT {q}" for q in range(n)] + [f"CX {q} {q + 1}" for q in range(n - 1)]
    lines += [f"H {q}\nT {q}" for q in range(n)] + ["M " + " ".join(map(str, range(n)))]
    return hip.Circuit("\n".join(lines) + "\n").compile(components="dense")


def test_moved_and_explicit_wide_group_shapes_share_a_batch(hip):
    """A width-12 and a width-11 program moved from the wide tier into the
    group buckets, a width-13 program asked to run at 512 threads per shot
    and a width-10 program batched together: every item's rows equal its
    own run's, and the buckets split by width class (up to 10, 11-12, 13)."""
    moved = dense_chain(hip, 12)
    moved11 = dense_chain(hip, 11)
    text13 = open(__file__.replace("test_gpu_batch.py", "golden/circuits/wide_compact.stim")).read()
    wide13 = hip.Circuit(text13).compile(tier="group", group_threads=512, components="dense")
    narrow = hip.Circuit(W.config0_circuit()).compile()
    assert moved.tier == "wide" and moved.max_local_k == 12
    assert moved11.tier == "wide" and moved11.max_local_k == 11
    assert wide13.tier == "group" and wide13.max_local_k == 13 and narrow.tier == "group"
    samplers, shots, seeds = [moved, wide13, narrow, moved11], [2000, 300, 1500, 900], [31, 32, 33, 34]
    outs = [buffers(s, k) for s, k in zip(samplers, shots)]
    b = hip.Batch(samplers, shots, seeds, outs, autotune=True, autotune_shots=256)
    b.launch(0)
    b.wait()
    info = b.info()
    print(info, flush=True)
    assert info["wide_items"] == 0
    # width classes: C0's width-10 bucket keeps its own shape; the width-11
    # and width-12 items need at least 128 and 256 threads per shot
    assert len(info["buckets"]) >= 3, info
    for s, k, sd, out in zip(samplers, shots, seeds, outs):
        ref = buffers(s, k)
        s.sample_device(k, sd, ref)
        torch.cuda.synchronize()
        for a, r in zip(out, ref):
            assert torch.equal(a, r), s.max_local_k
    b.close()


def test_wide_items_in_two_sequences_equal_single_runs(hip):
    """Wide items of a batch are split over two captured sequences on two
    streams (by estimated bytes), each with its own workspace; every item's
    rows and counts equal its single-program run, also with FP32 amplitude
    storage."""
    batch = W.config2_batch()
    for precision in ("fp64", "fp32"):
        samplers = [hip.Circuit(batch[i][0]).compile(tier="wide", precision=precision) for i in (3, 9, 27, 212, 250)]
        shots = [700, 1300, 500, 900, 1100]
        seeds = [5 + i for i in range(len(samplers))]
        outs = [buffers(s, n) for s, n in zip(samplers, shots)]
        b = hip.Batch(samplers, shots, seeds, outs, autotune=False)
        assert b.info()["wide_items"] == len(samplers)
        b.launch(3)
        b.wait()
        for s, n, sd, out in zip(samplers, shots, seeds, outs):
            ref = buffers(s, n)
            s.sample_device(n, sd + 3, ref)
            torch.cuda.synchronize()
            for a, r in zip(out, ref):
                assert torch.equal(a, r), (precision, s.max_active_qubits)
