"""Command line: sample a .ticit circuit on the hip backend.

Mirrors ticit's CLI (src/main.rs:16-67 arguments, 75-117 and 119-172 runs,
180-215 record files) with ``--backend hip``:

    python -m ticit_amd circuit.ticit -n 1000000 --seed 3 --records-out rows

prints ``key value`` lines (qubits, records, max_active_qubits, tier, shots,
discarded, accepted, logical_errors, keep_records, discard_rate,
logical_error_rate, sample_s) and, with ``--records-out P``, writes
``P.measurements.u8``, ``P.detectors.u8``, ``P.observables.u8`` (one byte
per bit, row-major), ``P.exp_vals.f64`` (little-endian) and ``P.meta``.
"""
from __future__ import annotations

import argparse
import math
import sys

import numpy as np


def _u8_list(text: str) -> list:
    return [int(v) for v in text.split(",") if v != ""]


def _usize_list(text: str) -> list:
    out = [int(v) for v in text.split(",") if v != ""]
    if any(v < 0 for v in out):
        raise argparse.ArgumentTypeError("values must be non-negative")
    return out


def _positive(text: str) -> int:
    v = int(text)
    if v < 1:
        raise argparse.ArgumentTypeError(f"{v} is not in 1..")
    return v


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(prog="python -m ticit_amd",
                                description="Sample a .ticit circuit with the hip backend")
    p.add_argument("circuit", help="circuit to sample")
    p.add_argument("-n", "--shots", type=_positive, default=1000, help="number of attempted shots")
    p.add_argument("--seed", type=int, default=1, help="seed for deterministic sampling")
    p.add_argument("--backend", default="hip", help="sampling backend (this package provides hip)")
    p.add_argument("-j", "--threads", type=_positive, default=1,
                   help="sampler threads (CPU backend option; the hip backend ignores it)")
    p.add_argument("--chunk-shots", type=_positive, default=1048576,
                   help="shots per launch group (the hip backend sizes its own launches)")
    p.add_argument("--postselection-mask", type=_u8_list, default=[],
                   help="flat detector postselection flags, separated by commas")
    p.add_argument("--normalize-syndromes", action="store_true",
                   help="XOR detector and observable outcomes with a noiseless reference sample")
    rec = p.add_mutually_exclusive_group()
    rec.add_argument("--records-out", default=None,
                     help="write per-shot measurement, detector, observable and expectation rows "
                          "using this path prefix")
    rec.add_argument("--count-only", action="store_true",
                     help="skip per-shot record retention and return aggregate counters only")
    p.add_argument("--exact-k", type=_usize_list, default=[],
                   help="condition every ordinary Bernoulli source on exactly this many faults")
    return p


def parse_args(argv=None) -> argparse.Namespace:
    p = build_parser()
    args = p.parse_args(argv)
    if args.exact_k and args.records_out is None:
        p.error("--exact-k requires --records-out")
    return args


def rate(value: float) -> str:
    """Rust's f64 Display: shortest round-trip digits, positional, no '.0'."""
    if math.isnan(value):
        return "nan"
    return np.format_float_positional(value, trim="-")


def write_records(prefix: str, sampler, result) -> None:
    """Record files in the layout of the reference's write_records."""
    with open(prefix + ".measurements.u8", "wb") as f:
        f.write(np.ascontiguousarray(result.measurements, dtype=np.uint8).tobytes())
    with open(prefix + ".detectors.u8", "wb") as f:
        f.write(np.ascontiguousarray(result.detectors, dtype=np.uint8).tobytes())
    with open(prefix + ".observables.u8", "wb") as f:
        f.write(np.ascontiguousarray(result.observables, dtype=np.uint8).tobytes())
    with open(prefix + ".exp_vals.f64", "wb") as f:
        f.write(np.ascontiguousarray(result.exp_vals, dtype="<f8").tobytes())
    rows = int(result.measurements.shape[0])
    with open(prefix + ".meta", "w") as f:
        f.write(f"rows {rows}\nmeasurements {sampler.num_measurements}\ndetectors {sampler.num_detectors}\n"
                f"observables {sampler.num_observables}\nexpectations {sampler.num_expectation_values}\n"
                "layout row-major\nf64 little-endian\n")


def main(argv=None) -> int:
    args = parse_args(argv)
    if args.backend != "hip":
        print(f"error: backend {args.backend!r} is not provided by ticit_amd; use --backend hip", file=sys.stderr)
        return 2
    if args.exact_k:
        print("error: --exact-k is not supported by the hip backend", file=sys.stderr)
        return 2
    from . import Circuit

    circuit = Circuit.from_file(args.circuit)
    sampler = circuit.compile(args.postselection_mask or None, normalize_syndromes=args.normalize_syndromes)
    keep = not args.count_only
    result = sampler.sample(args.shots, seed=args.seed, bit_packed=False, keep_records=keep)
    if args.records_out is not None:
        write_records(args.records_out, sampler, result)
    discard_rate = result.discarded / result.shots if result.shots else float("nan")
    print(f"qubits {sampler.num_qubits}")
    print(f"records {sampler.num_measurements}")
    print(f"max_active_qubits {sampler.max_active_qubits}")
    print(f"tier {sampler.tier}")
    print(f"shots {result.shots}")
    print(f"discarded {result.discarded}")
    print(f"accepted {result.accepted}")
    print(f"logical_errors {result.logical_errors}")
    print(f"keep_records {'true' if keep else 'false'}")
    print(f"discard_rate {rate(discard_rate)}")
    print(f"logical_error_rate {rate(result.logical_error_rate)}")
    print(f"sample_s {result.sample_s}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
