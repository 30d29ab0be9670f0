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
every host core this process may run on (at most 16).

Unless --headline-only is given, the same JSON line carries a "configs"
object with the other workloads: C0 (random Clifford+T, group tier), C2
(256 circuits: one-at-a-time versus one batched graph launch), C3 (dense
stress: FP64 versus FP32 amplitudes, probability error against the FP64
oracle), C4 (held-out planner batch: heuristic versus learned layouts,
one-at-a-time versus batched) and P (a cultivation circuit with every
detector postselected and normalized syndromes), each with a bounded
CPU-oracle sample.
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
    ap.add_argument("--headline-only", action="store_true",
                    help="only the surface-code line, without the other configurations")
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
            ones.data_ptr(), None, ctypes_stream(stream)))

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
    if not args.headline_only:
        line["configs"] = other_configs(ticit_amd, w, torch, np, stream, use_cpu=not args.no_cpu)
    print(json.dumps(line))
    return 0


def _cores() -> int:
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


def _device_buffers(torch, s, n):
    row = lambda c: (c + 7) // 8
    mk = lambda c: torch.empty((n, max(1, row(c))), dtype=torch.uint8, device="cuda:0")
    return (mk(s.num_measurements), mk(s.num_detectors), mk(s.num_observables),
            torch.zeros(max(1, s.num_observables), dtype=torch.int64, device="cuda:0"))


def _time(torch, stream, fn, reps=2):
    """Best wall time of `reps` runs of fn(rep) after one warm run, with the
    device drained before and after each."""
    fn(-1)
    torch.cuda.synchronize()
    best = float("inf")
    for r in range(reps):
        t = time.perf_counter()
        fn(r)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    return best


def other_configs(ticit_amd, w, torch, np, stream, use_cpu=True) -> dict:
    """The other benchmark configurations: C0 (random Clifford+T, group
    tier), C2 (256 circuits, plain versus one batched launch), C3 (dense
    stress, FP64 versus FP32 amplitudes, error against FP64), C4 (held-out
    planner batch: heuristic versus learned layouts, plain versus batched),
    P (postselected cultivation circuit, survivors compacted on the device)."""
    from oracle import oracle as O
    out = {}
    cores = _cores()
    st = stream.cuda_stream

    # C0
    c0 = ticit_amd.Circuit(w.config0_circuit())
    s0 = c0.compile()
    n0 = 100_000
    b0 = _device_buffers(torch, s0, n0)
    t = _time(torch, stream, lambda r: s0.sample_device(n0, 10 + r, b0, stream=st), reps=5)
    out["C0"] = {"workload": "random Clifford+T, 16 qubits, 40 layers, 12 T, M all; 1e5 shots, FP64",
                 "tier": s0.tier, "shots_per_s": n0 / t, "ms": t * 1e3}
    if use_cpu:
        r = O.sample(c0.blob, 20000, 3, threads=cores, fast=True)
        out["C0"]["cpu_baseline"] = {"value": 20000 / r["sample_s"], "unit": "shots/s", "cores": cores,
                                     "kind": "port", "sample": "20000 shots of C0"}

    # C2
    batch = w.config2_batch()
    samplers = [ticit_amd.Circuit(tx).compile() for tx, _ in batch]
    shots = [n for _, n in batch]
    seeds = [1 + i for i in range(len(batch))]
    bufs = [_device_buffers(torch, s, n) for s, n in zip(samplers, shots)]

    def plain(r):
        for s, n, sd, b in zip(samplers, shots, seeds, bufs):
            s.sample_device(n, sd + 100 * (r + 2), b, stream=st)

    tp = _time(torch, stream, plain, reps=3)
    bh = ticit_amd.Batch(samplers, shots, seeds, bufs, autotune=True, autotune_shots=4096)
    tbh = _time(torch, stream, lambda r: bh.launch(100 * (r + 2), st), reps=3)
    bh.close()
    learned = [ticit_amd.CompiledSampler(s.blob, planner="learned") for s in samplers]
    t0 = time.perf_counter()
    bt = ticit_amd.Batch(learned, shots, seeds, bufs, autotune=True, autotune_shots=4096)
    setup = time.perf_counter() - t0
    tb = _time(torch, stream, lambda r: bt.launch(100 * (r + 2), st), reps=3)
    info = bt.info()
    total = sum(shots)
    out["C2"] = {"workload": "256 random Clifford+T circuits, 8-40 qubits, 4-20 T, DEPOLARIZE1 5e-4, "
                             "log-uniform 1e4-1e6 shots",
                 "shots": total, "plain_shots_per_s": total / tp,
                 "batched_heuristic_shots_per_s": total / tbh,
                 "batched_learned_shots_per_s": total / tb,
                 "batched_learned_over_plain": tp / tb, "batched_heuristic_over_plain": tp / tbh,
                 "plain_s": tp, "batched_learned_s": tb, "batch_setup_s": setup,
                 "timing": "each leg: one warm run, then the best of 3 timed runs",
                 "components_chosen": {"heuristic": int(sum(s.components for s in samplers)),
                                       "learned": int(sum(s.components for s in learned))},
                 "tiers": {"lane": info["lane_items"], "wide": info["wide_items"],
                           "group": info["items"] - info["lane_items"] - info["wide_items"]},
                 "group_threads_per_shot": [bk["threads_per_shot"] for bk in info["buckets"]]}
    bt.close()
    if use_cpu:
        sel = list(range(0, 256, 16))
        n_cpu, t_cpu = 0, 0.0
        for i in sel:
            n = min(shots[i], 4000)
            r = O.sample(samplers[i].blob, n, 5, threads=cores, fast=True)
            n_cpu += n
            t_cpu += r["sample_s"]
        out["C2"]["cpu_baseline"] = {"value": n_cpu / t_cpu, "unit": "shots/s", "cores": cores, "kind": "port",
                                     "sample": f"circuits {sel[0]}..{sel[-1]} step 16, up to 4000 shots each"}
    del bufs

    # C3
    c3 = ticit_amd.Circuit(w.config3_circuit())
    n3 = 100_000
    res = {}
    for prec in ("fp64", "fp32"):
        s3 = c3.compile(precision=prec)
        b3 = _device_buffers(torch, s3, n3)
        t = _time(torch, stream, lambda r: s3.sample_device(n3, 20 + r, b3, stream=st), reps=1)
        res[prec] = (t, s3)
        del b3
    t64, s64 = res["fp64"]
    t32, s32 = res["fp32"]
    q = 8
    rng = np.random.default_rng(3)
    uni = rng.random((s64.num_measure_instructions, q))
    o = O.sample_injected(c3.blob, q, None, uni, None, want_probs=True)
    g32 = s32.sample_injected(q, None, uni, None, want_probs=True)
    g64 = s64.sample_injected(q, None, uni, None, want_probs=True)
    O.set_exact_sums(True)
    try:
        ox = O.sample_injected(c3.blob, q, None, uni, None, want_probs=True)
    finally:
        O.set_exact_sums(False)
    out["C3"] = {"workload": "30 qubits, 60 layers, 24 T, M on 8 qubits after layers 15/30/45, final M; 1e5 shots",
                 "layout": "components" if s64.components else "dense", "max_local_k": s64.max_local_k,
                 "fp64_shots_per_s": n3 / t64, "fp32_shots_per_s": n3 / t32, "fp32_over_fp64": t64 / t32,
                 "fp64_algorithmic_GBps": s64.algorithmic_bytes_per_shot * n3 / t64 / 1e9,
                 "fp32_algorithmic_GBps": s32.algorithmic_bytes_per_shot * n3 / t32 / 1e9,
                 "hbm_frac_fp64": s64.algorithmic_bytes_per_shot * n3 / t64 / 1e9 / HBM_PEAK_GBS,
                 "max_prob_error_fp64": float(np.max(np.abs(g64.probabilities - o["probs"]))),
                 "max_prob_error_fp32": float(np.max(np.abs(g32.probabilities - o["probs"]))),
                 "max_prob_error_fp64_vs_compensated": float(np.max(np.abs(g64.probabilities - ox["probs"]))),
                 "max_prob_error_oracle_scalar_vs_compensated": float(np.max(np.abs(o["probs"] - ox["probs"]))),
                 "error_sample": f"{q} shots, injected uniforms, against the FP64 oracle"}
    if use_cpu:
        r = O.sample(c3.blob, 32, 4, threads=cores, fast=True)
        out["C3"]["cpu_baseline"] = {"value": 32 / r["sample_s"], "unit": "shots/s", "cores": cores,
                                     "kind": "port", "sample": "32 shots of C3"}

    # C4
    from ticit_amd import planner as PL
    m = PL.load_model()
    ho = [int(i) for i in m["ho_idx"]]
    blobs = [ticit_amd.plan_text(w.planner_corpus_circuit(i)) for i in ho]
    n4 = 20_000
    heur = [ticit_amd.CompiledSampler(b, planner="heuristic") for b in blobs]
    lear = [ticit_amd.CompiledSampler(b, planner="learned") for b in blobs]
    bufs4 = [_device_buffers(torch, s, n4) for s in heur]

    def run_each(samps):
        return lambda r: [s.sample_device(n4, 7 + i + 1000 * (r + 2), b, stream=st)
                          for i, (s, b) in enumerate(zip(samps, bufs4))]

    th = _time(torch, stream, run_each(heur), reps=3)
    tl = _time(torch, stream, run_each(lear), reps=3)
    b4 = ticit_amd.Batch(lear, [n4] * len(lear), [7 + i for i in range(len(lear))], bufs4, autotune=True,
                         autotune_shots=4096)
    tbl = _time(torch, stream, lambda r: b4.launch(1000 * (r + 2), st), reps=3)
    b4.close()
    ev = PL.evaluate(m["ho_p_components"], m["ho_heuristic"], m["ho_td"], m["ho_tc"], float(m["margin"]))
    total4 = n4 * len(blobs)
    out["C4"] = {"workload": f"{len(blobs)} held-out planner-corpus programs (indices >= 100000), {n4} shots each",
                 "plain_heuristic_shots_per_s": total4 / th, "plain_learned_shots_per_s": total4 / tl,
                 "batched_learned_shots_per_s": total4 / tbl, "learned_over_heuristic": th / tl,
                 "batched_learned_over_plain_heuristic": th / tbl,
                 "components_chosen": {"heuristic": int(sum(s.components for s in heur)),
                                       "learned": int(sum(s.components for s in lear))},
                 "learned_fallbacks": int(sum(s.planner_source != "learned" for s in lear)),
                 "regret_learned": ev["regret_learned"], "regret_heuristic": ev["regret_heuristic"],
                 "accuracy_learned": ev["accuracy_learned"], "accuracy_heuristic": ev["accuracy_heuristic"],
                 "train_programs": int(m["n_train"])}
    if use_cpu:
        sel = list(range(0, len(blobs), 32))
        n_cpu, t_cpu = 0, 0.0
        for i in sel:
            r = O.sample(blobs[i], 2000, 5, threads=cores, fast=True)
            n_cpu += 2000
            t_cpu += r["sample_s"]
        out["C4"]["cpu_baseline"] = {"value": n_cpu / t_cpu, "unit": "shots/s", "cores": cores, "kind": "port",
                                     "sample": f"held-out programs {sel[0]}..{sel[-1]} step 32, 2000 shots each"}

    # Postselection: the cultivation fixture as the reference's benchmark
    # driver runs it (normalize_syndromes, a postselection mask), every
    # detector postselected; survivors compacted on the device.
    path = os.path.join(ROOT, "tests", "golden", "circuits", "msc_d3_inject_cultivate_p1e-3.stim")
    with open(path) as f:
        cp = ticit_amd.Circuit(f.read())
    mask = [1] * cp.num_detectors
    sp = cp.compile(mask, normalize_syndromes=True)
    npost = 1_000_000
    bp = _device_buffers(torch, sp, npost)
    acc = torch.zeros(1, dtype=torch.int64, device="cuda:0")
    t = _time(torch, stream, lambda r: sp.sample_device(npost, 20 + r, bp, stream=st, accepted=acc), reps=5)
    torch.cuda.synchronize()
    kept = int(acc.item())
    out["P"] = {"workload": "magic-state cultivation d=3 (injection + cultivation, p=1e-3), every detector "
                            "postselected, normalized syndromes; 1e6 shots, survivors compacted on the device",
                "tier": sp.tier, "shots_per_s": npost / t, "ms": t * 1e3, "accepted_fraction": kept / npost}
    if use_cpu:
        n_cpu = 20000
        O.lib(fast=True)  # loaded before the expected bits are set in both builds
        O.set_expected(*O.reference_sample(sp.blob))
        try:
            r = O.sample(sp.blob, n_cpu, 3, batch_size=1, threads=cores, fast=True)
        finally:
            O.set_expected(None, None)
        out["P"]["cpu_baseline"] = {"value": n_cpu / r["sample_s"], "unit": "shots/s", "cores": cores,
                                    "kind": "port", "sample": f"{n_cpu} shots of the same postselected program, "
                                                              "one shot per block"}
    return out


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
