"""Benchmark of the hip sampling backend on one MI355X.

Workload: rotated surface-code Z-memory, distance 5, 5 rounds, DEPOLARIZE1/2
and X_ERROR at p = 1e-3, two T gates per round (10 total), 120 detectors and
one logical observable, 1e6 shots per step, bit-packed rows (35 bytes per
shot) written to device memory. One step = one sampling call over the 1e6
shots; `value` is shots per second over the K timed steps.

Roofline: the dominant kernel is the resident lane-tier sampler. Its
algorithmic HBM bytes per shot are the packed output rows
(ceil(145/8) + ceil(120/8) + ceil(1/8) = 35 B): every amplitude, symbol and
record stays in LDS. achieved = 35 B * shots / kernel time, measured with HIP
events on the stream the kernel runs on; peak = 8 TB/s HBM3E.

CPU baseline: the oracle restatement of ticit's CPU sampler (reference
chunking and batch defaults) on a bounded sample of the same circuit, using
every host core this process may run on.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "lane_tier_surface_d5_traffic.json")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shots", type=int, default=1_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              "(shots shard by chunk index across GPUs; see the design notes)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import ticit_amd
    from ticit_amd import workloads as w

    text = w.config1_circuit()
    circuit = ticit_amd.Circuit(text)
    sampler = circuit.compile()
    shots = args.shots
    lib = ticit_amd._capi.load()
    dev = torch.device("cuda:0")
    row = lambda n: (n + 7) // 8
    meas = torch.empty((shots, row(sampler.num_measurements)), dtype=torch.uint8, device=dev)
    det = torch.empty((shots, row(sampler.num_detectors)), dtype=torch.uint8, device=dev)
    obs = torch.empty((shots, row(sampler.num_observables)), dtype=torch.uint8, device=dev)
    ones = torch.zeros(max(1, sampler.num_observables), dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    def step(seed: int) -> None:
        ticit_amd._capi.check(lib.ticit_hip_sample_device(
            sampler.handle, shots, seed, 0, 1, 1, meas.data_ptr(), det.data_ptr(), obs.data_ptr(),
            ones.data_ptr(), ctypes_stream(stream)))

    for i in range(args.warmup):
        step(1000 + i)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step(i + 1)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    avg_kernel_s = float(np.mean(kernel_ms)) * 1e-3
    value = shots * args.steps / wall

    bytes_per_shot = row(sampler.num_measurements) + row(sampler.num_detectors) + row(sampler.num_observables)
    achieved = bytes_per_shot * shots / avg_kernel_s / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            prof = json.load(f)
        if prof.get("shots") == shots:
            traffic = prof.get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(circuit.blob, args.cpu_seconds)

    line = {
        "metric": "shots/s per circuit on 1 MI355X + % HBM roofline",
        "value": value,
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "rotated surface-code Z-memory d=5, 5 rounds, p=1e-3 DEPOLARIZE1/2 + X_ERROR, "
                        "10 T gates, 1e6 shots, bit-packed rows in device memory",
            "shots_per_step": shots,
            "max_active_qubits": sampler.max_active_qubits,
            "tier": sampler.tier,
            "kernel_ms_per_step": avg_kernel_s * 1e3,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


def ctypes_stream(stream) -> int:
    return int(stream.cuda_stream)


def cpu_baseline(blob, seconds: float) -> dict:
    """Times the CPU oracle (reference defaults: 2048-shot chunks, default
    batch) on a bounded sample sized to about `seconds` of work."""
    from oracle import oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    probe = O.sample(blob, 20480, 1, threads=cores, fast=True)
    rate = 20480 / max(probe["sample_s"], 1e-6)
    shots = int(min(max(rate * seconds, 20480), 20_000_000))
    r = O.sample(blob, shots, 2, threads=cores, fast=True)
    return {
        "value": shots / r["sample_s"],
        "unit": "shots/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{shots} shots of the same circuit, seed 2, bit-packed rows kept, "
                  f"{cores} worker threads, chunk 2048 / default batch",
    }


if __name__ == "__main__":
    sys.exit(main())
