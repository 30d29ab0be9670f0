"""Traffic probe for rocprofv3 counter runs.

``python -m ticit_amd.profiling`` runs one device copy of a known size
(256 MiB read, 256 MiB written) and then two surface-code sampling launches
(1e6 shots each, bit-packed rows into device memory). Profiled with
``rocprofv3 --pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE``, the copy calibrates
each counter and the sampling launches give the kernel's HBM bytes.
"""
from __future__ import annotations

import sys


def main() -> int:
    import torch

    import ticit_amd
    from ticit_amd import workloads as w

    n = 256 << 20
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    s = ticit_amd.Circuit(w.config1_circuit()).compile()
    shots = 1_000_000
    dev = "cuda"
    meas = torch.empty((shots, 19), dtype=torch.uint8, device=dev)
    det = torch.empty((shots, 15), dtype=torch.uint8, device=dev)
    obs = torch.empty((shots, 1), dtype=torch.uint8, device=dev)
    ones = torch.zeros(1, dtype=torch.int64, device=dev)
    lib = ticit_amd._capi.load()
    for seed in (1, 2):
        ticit_amd._capi.check(lib.ticit_hip_sample_device(
            s.handle, shots, seed, 0, 1, 1, meas.data_ptr(), det.data_ptr(), obs.data_ptr(),
            ones.data_ptr(), None))
    torch.cuda.synchronize()
    print("traffic probe done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
