"""Probe workloads for rocprofv3 runs.

``python -m ticit_amd.profiling [c1|c0|c3|c2|p]`` runs one device copy of a
known size (256 MiB read, 256 MiB written) and then the workload's sampling
launches into device memory:

  * c1 (default): surface code d=5, 5 rounds, two launches of 1e6 shots;
  * c0: random Clifford+T n=16, 12 T (group tier), two launches of 1e5 shots;
  * c3: dense stress n=30, 24 T (wide tier, tensor components, FP64), one
    warm launch and one measured launch of 4096 shots;
  * c2: the 256-circuit batch as one batched graph launch (learned layouts);
  * p: the d=3 cultivation fixture, every detector postselected, normalized,
    two launches of 1e6 shots with the survivors compacted on the device.

Profiled with ``rocprofv3 --pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE`` (one
counter per run), the copy calibrates each counter and the sampling
launches give the kernels' HBM bytes; with ``--kernel-trace --stats`` the
per-kernel times.
"""
from __future__ import annotations

import sys


def _buffers(torch, s, shots):
    row = lambda n: max(1, (n + 7) // 8)
    return (torch.empty((shots, row(s.num_measurements)), dtype=torch.uint8, device="cuda"),
            torch.empty((shots, row(s.num_detectors)), dtype=torch.uint8, device="cuda"),
            torch.empty((shots, row(s.num_observables)), dtype=torch.uint8, device="cuda"),
            torch.zeros(max(1, s.num_observables), dtype=torch.int64, device="cuda"))


def main(argv) -> int:
    import torch

    import ticit_amd
    from ticit_amd import workloads as w

    which = argv[0] if argv else "c1"
    n = 256 << 20
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    if which in ("c1", "c0", "c3"):
        text = {"c1": w.config1_circuit, "c0": w.config0_circuit, "c3": w.config3_circuit}[which]()
        s = ticit_amd.Circuit(text).compile()
        shots = {"c1": 1_000_000, "c0": 100_000, "c3": 4096}[which]
        out = _buffers(torch, s, shots)
        for seed in (1, 2):
            s.sample_device(shots, seed, out)
        torch.cuda.synchronize()
    elif which == "c2":
        batch = w.config2_batch()
        samplers = [ticit_amd.CompiledSampler(ticit_amd.plan_text(t), planner="learned") for t, _ in batch]
        shots = [n for _, n in batch]
        outs = [_buffers(torch, s, n) for s, n in zip(samplers, shots)]
        bt = ticit_amd.Batch(samplers, shots, [1 + i for i in range(len(batch))], outs, autotune=True,
                             autotune_shots=4096)
        bt.launch(0)
        bt.wait()
        torch.cuda.synchronize()
    elif which == "p":
        # the cultivation fixture, every detector postselected, normalized;
        # survivors compacted on the device
        import os
        path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden",
                            "circuits", "msc_d3_inject_cultivate_p1e-3.stim")
        with open(path) as f:
            c = ticit_amd.Circuit(f.read())
        s = c.compile([1] * c.num_detectors, normalize_syndromes=True)
        shots = 1_000_000
        out = _buffers(torch, s, shots)
        acc = torch.zeros(1, dtype=torch.int64, device="cuda")
        for seed in (1, 2):
            s.sample_device(shots, seed, out, accepted=acc)
        torch.cuda.synchronize()
        print(f"accepted {int(acc.item())} of {shots}", flush=True)
    else:
        print(f"unknown probe {which!r}", file=sys.stderr)
        return 2
    print(f"probe {which} done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
