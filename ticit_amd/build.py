"""In-tree build of the hip sampling backend: ``libticit_hip.so`` for gfx950.

Host C++ (front end, device lowering, C ABI) and the HIP kernels are compiled
by hipcc into one shared library under ``ticit_amd/_lib``. Device code is
built without FP contraction (the dense kernels' evaluation order is part of
the parity contract) and with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "libticit_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = [
    "csrc/host/frontend.cpp",
    "csrc/host/program.cpp",
    "csrc/host/component_plan.cpp",
    "csrc/host/pinning.cpp",
    "csrc/host/device_program.cpp",
    "csrc/host/capi.cpp",
    "csrc/host/host_copy.cpp",
    "csrc/host/batch.cpp",
    "csrc/host/planner.cpp",
    "csrc/kernels/sampler.hip",
    "csrc/kernels/group_tier.hip",
    "csrc/kernels/wide_tier.hip",
    "csrc/kernels/compact.hip",
    "csrc/kernels/static_scratch.hip",
    "csrc/kernels/planner_mlp.hip",
]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(HERE, src)
        obj = os.path.join(LIB_DIR, os.path.basename(src) + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], "-c", *lang, path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
        if p.returncode != 0:
            sys.stderr.write("\n".join(lines) + "\n")
            raise RuntimeError(f"hipcc failed: {' '.join(cmd)}")
        if verbose and lines:
            print("\n".join(lines))
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", LIB + ".tmp"]
    subprocess.run(link, check=True)
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
