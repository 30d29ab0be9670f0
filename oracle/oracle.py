"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module, always as the checker. It loads ``oracle/_build/liboracle.so``
(parity build) or ``liboracle_fast.so`` (timing build) produced by
``oracle/Makefile``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def build_native() -> str:
    """Builds the -march=native timing build for this host (seconds of g++);
    returns its path."""
    subprocess.run(["make", "-s", "-C", HERE, "native"], check=True)
    return os.path.join(BUILD, "native", "liboracle_native.so")


def _load(name: str) -> ctypes.CDLL:
    path = os.path.join(BUILD, name)
    if name == "native/liboracle_native.so":
        path = build_native()
    elif not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    lib.oracle_last_error.restype = ctypes.c_char_p
    lib.oracle_info.argtypes = [_u64p, ctypes.c_size_t, _i64p]
    lib.oracle_sample.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _u8p, _u8p, _u8p, _u64p, _u64p, _f64p, _f64p]
    lib.oracle_sample_injected.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_uint64, _u64p, _f64p,
                                           _u64p, ctypes.c_int, ctypes.c_int, _u8p, _u8p, _u8p,
                                           _u64p, _u64p, _f64p, _f64p, _f64p,
                                           ctypes.POINTER(ctypes.c_int), _f64p]
    lib.oracle_presample.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64, _u64p]
    lib.oracle_splitmix_next.argtypes = [_u64p]
    lib.oracle_splitmix_next.restype = ctypes.c_uint64
    lib.oracle_rand_float.argtypes = [_u64p]
    lib.oracle_rand_float.restype = ctypes.c_double
    lib.oracle_block_seed.argtypes = [ctypes.c_uint64] * 3
    lib.oracle_block_seed.restype = ctypes.c_uint64
    lib.oracle_categorical_row.argtypes = [_u64p, _f64p, ctypes.c_size_t]
    lib.oracle_categorical_row.restype = ctypes.c_uint64
    lib.oracle_bernoulli.argtypes = [_u64p, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_insert_zero_bit.argtypes = [ctypes.c_size_t, ctypes.c_int]
    lib.oracle_insert_zero_bit.restype = ctypes.c_size_t
    lib.oracle_set_fp32_storage.argtypes = [ctypes.c_int]
    lib.oracle_set_component_mode.argtypes = [ctypes.c_int]
    lib.oracle_set_exact_sums.argtypes = [ctypes.c_int]
    lib.oracle_set_keep_records.argtypes = [ctypes.c_int]
    lib.oracle_set_forced.argtypes = [_i64p, ctypes.c_size_t]
    lib.oracle_set_expected.argtypes = [_u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t]
    lib.oracle_set_expected.restype = None
    lib.oracle_reference_sample.argtypes = [_u64p, ctypes.c_size_t, _u8p, _u8p]
    lib.oracle_component_plan.argtypes = [_u64p, ctypes.c_size_t, _i64p, _f64p, _i64p, _i64p, _i64p,
                                          ctypes.c_size_t]
    lib.oracle_default_batch.argtypes = [ctypes.c_int]
    lib.oracle_default_batch.restype = ctypes.c_int64
    return lib


_LIBS: dict[str, ctypes.CDLL] = {}


def lib(fast=False) -> ctypes.CDLL:
    """fast = False: the parity build; True: the portable timing build;
    "native": the timing build compiled for this host."""
    name = ("native/liboracle_native.so" if fast == "native" else
            "liboracle_fast.so" if fast else "liboracle.so")
    if name not in _LIBS:
        _LIBS[name] = _load(name)
    return _LIBS[name]


class OracleError(RuntimeError):
    pass


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _check(rc: int, fast: bool = False) -> None:
    if rc != 0:
        raise OracleError(lib(fast).oracle_last_error().decode())


def info(blob: np.ndarray) -> dict:
    out = np.zeros(11, dtype=np.int64)
    _check(lib().oracle_info(_ptr(blob, _u64p), blob.size, _ptr(out, _i64p)))
    keys = ["n", "initial_k", "max_k", "nsymbols", "nrecords", "ndetectors", "nobservables",
            "ninstr", "nmeasure", "ncoin", "nexpvals"]
    return dict(zip(keys, (int(v) for v in out)))


def _cols(n: int, packed: bool) -> int:
    return (n + 7) // 8 if packed else n


def sample(blob: np.ndarray, shots: int, seed: int, *, chunk_shots: int = 0, batch_size: int = 0,
           threads: int = 1, bit_packed: bool = True, observable: int = 0, fast=False,
           keep_records: bool = True) -> dict:
    """Reference-stream sampling; chunk/batch 0 pick the reference defaults.
    keep_records = False samples counts only (no rows)."""
    inf = info(blob)
    n = shots if keep_records else 0
    meas = np.zeros((n, _cols(inf["nrecords"], bit_packed)), dtype=np.uint8)
    det = np.zeros((n, _cols(inf["ndetectors"], bit_packed)), dtype=np.uint8)
    obs = np.zeros((n, _cols(inf["nobservables"], bit_packed)), dtype=np.uint8)
    counts = np.zeros(4, dtype=np.uint64)
    ones = np.zeros(max(1, inf["nobservables"]), dtype=np.uint64)
    timing = np.zeros(1, dtype=np.float64)
    ev = np.zeros((n, inf["nexpvals"]), dtype=np.float64)
    L = lib(fast)
    L.oracle_set_keep_records(int(keep_records))
    try:
        _check(L.oracle_sample(_ptr(blob, _u64p), blob.size, shots, seed, chunk_shots, batch_size,
                               threads, int(bit_packed), observable, _ptr(meas, _u8p), _ptr(det, _u8p),
                               _ptr(obs, _u8p), _ptr(counts, _u64p), _ptr(ones, _u64p),
                               _ptr(timing, _f64p), _ptr(ev, _f64p)), fast)
    finally:
        L.oracle_set_keep_records(1)
    rows = int(counts[2]) if keep_records else 0  # postselection keeps the accepted shots' rows
    return dict(measurements=meas[:rows], detectors=det[:rows], observables=obs[:rows], counts=counts,
                observable_ones=ones[:inf["nobservables"]], sample_s=float(timing[0]), exp_vals=ev[:rows])


def sample_per_shot(blob: np.ndarray, shots: int, seed: int, **kw) -> dict:
    """The per-shot stream contract: reference sampling with one-shot chunks
    and one-shot blocks, so shot s draws from block_seed(.., seed, s)."""
    return sample(blob, shots, seed, chunk_shots=1, batch_size=1, **kw)


def sample_injected(blob: np.ndarray, shots: int, exo: np.ndarray | None, uniforms: np.ndarray | None,
                    coins: np.ndarray | None, *, bit_packed: bool = True, observable: int = 0,
                    want_probs: bool = False, want_final_state: bool = False) -> dict:
    inf = info(blob)
    words = (shots + 63) // 64
    exo = np.ascontiguousarray(exo if exo is not None else np.zeros((inf["nsymbols"], words), np.uint64),
                               dtype=np.uint64)
    uniforms = np.ascontiguousarray(uniforms if uniforms is not None else np.zeros((inf["nmeasure"], shots)),
                                    dtype=np.float64)
    coins = np.ascontiguousarray(coins if coins is not None else np.zeros((inf["ncoin"], words), np.uint64),
                                 dtype=np.uint64)
    meas = np.zeros((shots, _cols(inf["nrecords"], bit_packed)), dtype=np.uint8)
    det = np.zeros((shots, _cols(inf["ndetectors"], bit_packed)), dtype=np.uint8)
    obs = np.zeros((shots, _cols(inf["nobservables"], bit_packed)), dtype=np.uint8)
    counts = np.zeros(4, dtype=np.uint64)
    ones = np.zeros(max(1, inf["nobservables"]), dtype=np.uint64)
    probs = np.zeros((inf["nmeasure"], shots)) if want_probs else None
    dim = 1 << inf["max_k"]
    fre = np.zeros(dim) if want_final_state else None
    fim = np.zeros(dim) if want_final_state else None
    fk = ctypes.c_int(0)
    ev = np.zeros((shots, inf["nexpvals"]), dtype=np.float64)
    _check(lib().oracle_sample_injected(_ptr(blob, _u64p), blob.size, shots, _ptr(exo, _u64p),
                                        _ptr(uniforms, _f64p), _ptr(coins, _u64p), int(bit_packed), observable,
                                        _ptr(meas, _u8p), _ptr(det, _u8p), _ptr(obs, _u8p), _ptr(counts, _u64p),
                                        _ptr(ones, _u64p), _ptr(probs, _f64p), _ptr(fre, _f64p), _ptr(fim, _f64p),
                                        ctypes.byref(fk), _ptr(ev, _f64p)))
    rows = int(counts[2])
    out = dict(measurements=meas[:rows], detectors=det[:rows], observables=obs[:rows], counts=counts,
               observable_ones=ones[:inf["nobservables"]], exp_vals=ev[:rows])
    if want_probs:
        out["probs"] = probs
    if want_final_state:
        d = 1 << fk.value
        out["final_state"] = fre[:d] + 1j * fim[:d]
    return out


def set_fp32_storage(on: bool) -> None:
    """Emulate FP32 amplitude storage in injected runs (precision studies)."""
    lib().oracle_set_fp32_storage(int(on))


def set_exact_sums(on: bool) -> None:
    """Compensated (Neumaier) probability sums in injected runs: a
    reference closer to exact than the reference sampler's scalar order."""
    lib().oracle_set_exact_sums(int(on))


def set_component_mode(mode: int) -> None:
    """-1: run as the blob's use_active_components flag says; 0: dense;
    1: tensor components."""
    for name, kind in (("liboracle.so", False), ("liboracle_fast.so", True), ("native/liboracle_native.so", "native")):
        if name in _LIBS or name != "native/liboracle_native.so":
            lib(kind).oracle_set_component_mode(int(mode))


STEP_KINDS = {0: "none", 1: "ignored_global_phase", 2: "rotation", 3: "promotion", 4: "measurement"}


def component_plan(blob: np.ndarray) -> dict:
    """The component plan of src/component_plan.rs:301-522 for a blob."""
    inf = info(blob)
    cap = max(64, inf["ninstr"] + 1, 4 * (inf["n"] + inf["ninstr"]))
    summary = np.zeros(9, dtype=np.int64)
    work = np.zeros(2)
    cmk = np.zeros(cap, dtype=np.int64)
    steps = np.zeros(6 * cap, dtype=np.int64)
    merges = np.zeros(cap, dtype=np.int64)
    _check(lib().oracle_component_plan(_ptr(blob, _u64p), blob.size, _ptr(summary, _i64p), _ptr(work, _f64p),
                                       _ptr(cmk, _i64p), _ptr(steps, _i64p), _ptr(merges, _i64p), cap))
    ncomp, nmerge, planned = int(summary[2]), int(summary[7]), bool(summary[8])
    rows = steps.reshape(cap, 6)[:inf["ninstr"] if planned else 0]
    return dict(selected=bool(summary[0]), initial_components=int(summary[1]), component_count=ncomp,
                dense_peak_dimension=int(summary[3]), component_peak_live_dimension=int(summary[4]),
                component_allocated_dimension=int(summary[5]), quantum_ops=int(summary[6]),
                planned=planned, estimated_dense_work=float(work[0]), estimated_component_work=float(work[1]),
                component_max_k=[int(v) for v in cmk[:ncomp]], merges=[int(v) for v in merges[:nmerge]],
                steps=[dict(kind=STEP_KINDS[int(r[0])], component=int(r[1]), merge_offset=int(r[2]),
                            merge_count=int(r[3]), pivot=int(r[4]), deactivate_after=bool(r[5])) for r in rows])


def presample(blob: np.ndarray, shots: int, stream_seed: int) -> np.ndarray:
    inf = info(blob)
    words = (shots + 63) // 64
    planes = np.zeros((max(inf["nsymbols"], 1), max(words, 1)), dtype=np.uint64)
    _check(lib().oracle_presample(_ptr(blob, _u64p), blob.size, shots, stream_seed, _ptr(planes, _u64p)))
    return planes[:inf["nsymbols"], :words]


def splitmix_next(state: int) -> tuple[int, int]:
    s = ctypes.c_uint64(state)
    v = lib().oracle_splitmix_next(ctypes.byref(s))
    return int(v), int(s.value)


def rand_float(state: int) -> tuple[float, int]:
    s = ctypes.c_uint64(state)
    v = lib().oracle_rand_float(ctypes.byref(s))
    return float(v), int(s.value)


def block_seed(base: int, seed: int, block: int) -> int:
    return int(lib().oracle_block_seed(base, seed, block))


def categorical_row(state: int, probs) -> tuple[int, int]:
    s = ctypes.c_uint64(state)
    p = np.ascontiguousarray(probs, dtype=np.float64)
    row = lib().oracle_categorical_row(ctypes.byref(s), _ptr(p, _f64p), p.size)
    return int(row), int(s.value)


def bernoulli(state: int, p: float) -> tuple[bool, int]:
    s = ctypes.c_uint64(state)
    out = ctypes.c_int(0)
    _check(lib().oracle_bernoulli(ctypes.byref(s), p, ctypes.byref(out)))
    return bool(out.value), int(s.value)


def insert_zero_bit(packed: int, bit: int) -> int:
    return int(lib().oracle_insert_zero_bit(packed, bit))


def default_batch(max_k: int) -> int:
    return int(lib().oracle_default_batch(max_k))


def set_pins(blob: np.ndarray | None, constraints=None) -> list:
    """Pins (records, value) constraints for the following runs on this
    process (oracle/pinning.py solves them); no constraints clears. Returns
    the forced branches as (instruction, constant, branch symbols)."""
    from oracle.pinning import plan_pinned
    forced = plan_pinned(blob, constraints) if constraints else []
    spec = []
    for instr, c, conds in forced:
        spec += [instr, int(c), len(conds), *conds]
    arr = np.ascontiguousarray(spec or [0], dtype=np.int64)
    for name, kind in (("liboracle.so", False), ("liboracle_fast.so", True), ("native/liboracle_native.so", "native")):
        if name in _LIBS or name == "liboracle.so":
            lib(kind).oracle_set_forced(_ptr(arr, _i64p), len(spec))
    return forced


def set_expected(detectors=None, observables=None) -> None:
    """Expected detector / observable bits for the following runs on this
    process (nonzero inverts the column); None or empty clears."""
    dl = [int(bool(v)) for v in (detectors if detectors is not None else [])]
    ol = [int(bool(v)) for v in (observables if observables is not None else [])]
    det = np.ascontiguousarray(dl or [0], dtype=np.uint8)
    obs = np.ascontiguousarray(ol or [0], dtype=np.uint8)
    nd, no = len(dl), len(ol)
    for name, kind in (("liboracle.so", False), ("liboracle_fast.so", True), ("native/liboracle_native.so", "native")):
        if name in _LIBS or name == "liboracle.so":
            lib(kind).oracle_set_expected(_ptr(det, _u8p), nd, _ptr(obs, _u8p), no)


def reference_sample(blob: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The noiseless reference sample: (detector bits, observable bits),
    obeying the pins set with set_pins."""
    inf = info(blob)
    det = np.zeros(max(1, inf["ndetectors"]), dtype=np.uint8)
    obs = np.zeros(max(1, inf["nobservables"]), dtype=np.uint8)
    _check(lib().oracle_reference_sample(_ptr(blob, _u64p), blob.size, _ptr(det, _u8p), _ptr(obs, _u8p)))
    return det[:inf["ndetectors"]], obs[:inf["nobservables"]]
