"""ticit_amd: the MI355X ``hip`` sampling backend for ticit's compiled
Clifford+T sampler.

The Python surface mirrors ticit's own binding (ticit_py/src/lib.rs:715-830,
tests in ticit_py/tests/test_api.py): ``Circuit(text).compile(...)`` returns a
compiled sampler whose ``sample(shots, seed, bit_packed)`` returns row-major
measurement/detector/observable records plus counts, with the same field
names. Everything below calls the C ABI of ``libticit_hip.so``
(include/ticit_hip.h); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import InvalidInputError, ParseError, TicitError
from .planner import planner_features

__all__ = ["Batch", "Circuit", "CompiledSampler", "InvalidInputError", "ParseError", "ReferenceSample", "SampleResult",
           "TicitError", "plan_features", "plan_text"]

TIERS = {"auto": 0, "lane": 1, "group": 2, "wide": 3}
# lane tier table placement (ticit_hip_options.lane_store)
LANE_STORES = {"auto": 0, "lds": 1, "bits_global": 2, "all_global": 3}
PRECISIONS = {"fp64": 0, "fp32": 1}
PLANNERS = {"heuristic": 0, "learned": 1}
COMPONENTS = {"planned": 0, "dense": 1, "components": 2}


def plan_text(text: str, postselection_mask=None) -> np.ndarray:
    """Parse and plan ``.ticit`` text into a plan blob (include/ticit_plan.h);
    ``postselection_mask`` (flat detector flags) postselects detectors, as
    SamplerOptions::postselection_mask does, unioned with DISCARD lines."""
    lib = _capi.load()
    data = text.encode()
    ptr = _capi._u64p()
    n = ctypes.c_size_t(0)
    if postselection_mask is None:
        _capi.check(lib.ticit_hip_plan_text(data, len(data), ctypes.byref(ptr), ctypes.byref(n)))
    else:
        mask = np.ascontiguousarray([1 if int(v) else 0 for v in postselection_mask], dtype=np.uint8)
        _capi.check(lib.ticit_hip_plan_text_postselected(data, len(data), mask.ctypes.data if mask.size else None,
                                                         mask.size, ctypes.byref(ptr), ctypes.byref(n)))
    try:
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy()
    finally:
        lib.ticit_hip_blob_free(ptr)


@dataclass
class SampleResult:
    """Mirrors ticit's SampleResult (src/sampler/prepared.rs:197-219)."""

    shots: int
    discarded: int
    accepted: int
    logical_errors: int
    bit_packed: bool
    measurements: np.ndarray
    detectors: np.ndarray
    observables: np.ndarray
    observable_ones: np.ndarray
    sample_s: float
    kernel_s: float
    probabilities: np.ndarray | None = None
    # [rows][num_expectation_values] EXP_VAL values (SampleResult::exp_vals)
    exp_vals: np.ndarray | None = None

    # ticit_py aliases
    @property
    def total_shots(self) -> int:
        return self.shots

    @property
    def passed_shots(self) -> int:
        return self.accepted

    @property
    def discards(self) -> int:
        return self.discarded

    @property
    def logical_error_rate(self) -> float:
        return self.logical_errors / self.accepted if self.accepted else float("nan")

    def __iter__(self):
        return iter((self.measurements, self.detectors, self.observables))


def _rows(ptr, rows: int, cols: int) -> np.ndarray:
    if rows == 0 or cols == 0 or not ptr:
        return np.zeros((rows, cols), dtype=np.uint8)
    return np.ctypeslib.as_array(ptr, shape=(rows * cols,)).copy().reshape(rows, cols)


def _convert(res: _capi.Result, keep_records: bool) -> SampleResult:
    packed = bool(res.bit_packed)
    cols = (lambda n: (n + 7) // 8) if packed else (lambda n: n)
    rows = int(res.record_rows) if keep_records else 0
    nobs = int(res.nobservables)
    ones = (np.ctypeslib.as_array(res.observable_ones, shape=(nobs,)).copy()
            if nobs and res.observable_ones else np.zeros(0, dtype=np.uint64))
    nexp = int(res.nexpvals)
    ev = (np.ctypeslib.as_array(res.exp_vals, shape=(rows * nexp,)).copy().reshape(rows, nexp)
          if rows and nexp and res.exp_vals else np.zeros((rows, nexp), dtype=np.float64))
    return SampleResult(
        shots=int(res.shots), discarded=int(res.discarded), accepted=int(res.accepted),
        logical_errors=int(res.logical_errors), bit_packed=packed,
        measurements=_rows(res.measurements, rows, cols(int(res.nmeasurements))),
        detectors=_rows(res.detectors, rows, cols(int(res.ndetectors))),
        observables=_rows(res.observables, rows, cols(nobs)),
        observable_ones=ones.astype(np.uint64), sample_s=float(res.sample_s), kernel_s=float(res.kernel_s),
        exp_vals=ev)


class ReferenceSample:
    """Circuit::reference_sample's result (src/sampler/prepared.rs:221-228):
    one bool per detector and per observable."""

    def __init__(self, detectors: list, observables: list):
        self.detectors = detectors
        self.observables = observables

    def __repr__(self) -> str:
        return f"ReferenceSample(detectors={self.detectors}, observables={self.observables})"


class CompiledSampler:
    """A compiled circuit bound to the current HIP device (Sampler,
    src/sampler/prepared.rs:467-478)."""

    backend = "hip"

    def __init__(self, blob: np.ndarray, observable: int = 0, precision: str = "fp64",
                 tier: str = "auto", planner: str = "heuristic", wide_batch: int = 0,
                 components: str = "planned", group_threads: int = 0, pin_measurements=None,
                 normalize_syndromes: bool = False, expected_detectors=None, expected_observables=None,
                 lane_store: str = "auto"):
        lib = _capi.load()
        self._lib = lib
        self.blob = np.ascontiguousarray(blob, dtype=np.uint64)
        if planner == "learned":
            from .planner import ensure_installed
            ensure_installed()
        opt = _capi.Options()
        opt.observable = observable
        opt.precision = PRECISIONS[precision]
        opt.tier = TIERS[tier]
        opt.planner_mode = PLANNERS[planner]
        opt.wide_batch = wide_batch
        opt.components = COMPONENTS[components]
        opt.group_threads = group_threads
        opt.lane_store = LANE_STORES[lane_store]
        handle = ctypes.c_void_p()
        reference = None
        if normalize_syndromes or expected_detectors is not None or expected_observables is not None:
            # SamplerOptions::normalize_syndromes / expected_detectors /
            # expected_observables (src/sampler/prepared.rs:111-116)
            det = np.ascontiguousarray([int(bool(v)) for v in (expected_detectors or [])], dtype=np.uint8)
            obs = np.ascontiguousarray([int(bool(v)) for v in (expected_observables or [])], dtype=np.uint8)
            reference = _capi.ReferenceBits()
            reference.detectors = det.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            reference.ndetectors = det.size
            reference.observables = obs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            reference.nobservables = obs.size
            reference.normalize_syndromes = int(bool(normalize_syndromes))
            self._reference_arrays = (det, obs)
        if pin_measurements or reference is not None:
            # SamplerOptions::pin_measurements: (records, value) pairs
            pin_measurements = pin_measurements or []
            recs = [np.ascontiguousarray([int(r) for r in records], dtype=np.uint64)
                    for records, _ in pin_measurements]
            pins = (_capi.Parity * max(1, len(recs)))()
            for j, (r, (_, value)) in enumerate(zip(recs, pin_measurements)):
                pins[j].records = r.ctypes.data_as(_capi._u64p)
                pins[j].nrecords = r.size
                pins[j].value = int(bool(value))
            _capi.check(lib.ticit_hip_program_create_ex(
                self.blob.ctypes.data_as(_capi._u64p), self.blob.size, ctypes.byref(opt), pins, len(recs),
                ctypes.byref(reference) if reference is not None else None, ctypes.byref(handle)))
        else:
            _capi.check(lib.ticit_hip_program_create(self.blob.ctypes.data_as(_capi._u64p), self.blob.size,
                                                     ctypes.byref(opt), ctypes.byref(handle)))
        self._handle = handle
        info = _capi.ProgramInfo()
        _capi.check(lib.ticit_hip_program_info_get(handle, ctypes.byref(info)))
        self.num_qubits = int(info.qubits)
        self.num_measurements = int(info.measurement_records)
        self.num_detectors = int(info.detectors)
        self.num_observables = int(info.observables)
        self.max_active_qubits = int(info.max_active_qubits)
        self.num_symbols = int(info.symbols)
        self.num_instructions = int(info.instructions)
        self.num_measure_instructions = int(info.measure_instructions)
        self.num_coin_instructions = int(info.coin_instructions)
        self.tier = {1: "lane", 2: "group", 3: "wide"}.get(int(info.tier), "auto")
        self.precision = precision
        self.wide_batch = int(info.wide_batch)
        self.output_bytes_per_shot = float(info.output_bytes_per_shot)
        self.state_bytes_per_shot = float(info.state_bytes_per_shot)
        self.algorithmic_bytes_per_shot = float(info.algorithmic_bytes_per_shot)
        self.components = bool(info.components)
        self.state_dim = int(info.state_dim)
        self.component_count = int(info.component_count)
        self.max_local_k = int(info.max_local_k)
        self.planner_source = {0: "heuristic", 1: "learned", 2: "caller"}[int(info.planner_source)]
        self.planner_p_components = float(info.planner_p_components)
        self.lds_bytes_per_block = int(info.lds_bytes_per_block)
        self.blocks_per_cu = int(info.blocks_per_cu)
        self.table_words = [int(v) for v in info.table_words]
        self.num_expectation_values = int(info.expectation_values)
        self.num_postselected_detectors = int(info.postselected_detectors)
        self.has_postselection = self.num_postselected_detectors > 0
        # lane tier: where a wave's tables live ("lds", "bits_global": bit
        # tables in device memory, "all_global": amplitudes there too)
        self.lane_store = {0: "lds", 1: "bits_global", 2: "all_global"}[int(info.lane_store)]
        self.lane_region_bytes = int(info.lane_region_bytes)

    def reference_sample(self) -> "ReferenceSample":
        """The noiseless reference sample of this program, obeying its pins
        (reference_sample_for_program, src/sampler/prepared.rs:355-405),
        computed on the device."""
        det = np.zeros(max(1, self.num_detectors), dtype=np.uint8)
        obs = np.zeros(max(1, self.num_observables), dtype=np.uint8)
        _capi.check(self._lib.ticit_hip_reference_sample(self._handle, det.ctypes.data, obs.ctypes.data))
        return ReferenceSample([bool(v) for v in det[:self.num_detectors]],
                               [bool(v) for v in obs[:self.num_observables]])

    def close(self) -> None:
        if getattr(self, "_handle", None):
            self._lib.ticit_hip_program_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._handle

    def sample(self, shots: int, seed: int | None = None, bit_packed: bool = False,
               keep_records: bool = True) -> SampleResult:
        """Sampler::sample_with_seed (src/sampler/prepared.rs:663-670)."""
        if seed is None:
            seed = int(np.random.SeedSequence().generate_state(1, np.uint64)[0])
        res = _capi.Result()
        _capi.check(self._lib.ticit_hip_sample(self._handle, shots, seed, int(bit_packed), int(keep_records),
                                               ctypes.byref(res)))
        try:
            return _convert(res, keep_records)
        finally:
            self._lib.ticit_hip_result_free(ctypes.byref(res))

    sample_with_seed = sample

    def sample_device(self, shots: int, seed: int, outputs, first_shot: int = 0, bit_packed: bool = True,
                      keep_records: bool = True, stream: int = 0, accepted=None) -> None:
        """Sampling into caller device buffers (ticit_hip_sample_device):
        ``outputs`` = (measurements, detectors, observables, observable_ones
        [, exp_vals]) objects with ``data_ptr()``; enqueued on ``stream``,
        not synchronized. With ``accepted`` (a device uint64 / int64 word)
        the call is ticit_hip_sample_device_postselected: the survivors' rows
        come first, in shot order, and their number lands in ``accepted``."""
        ptr = lambda t: ctypes.c_void_p(int(t.data_ptr())) if t is not None else None
        ev = outputs[4] if len(outputs) > 4 else None
        if accepted is not None:
            _capi.check(self._lib.ticit_hip_sample_device_postselected(
                self._handle, shots, seed, first_shot, int(bit_packed), int(keep_records), ptr(outputs[0]),
                ptr(outputs[1]), ptr(outputs[2]), ptr(outputs[3]), ptr(ev), ptr(accepted),
                ctypes.c_void_p(stream or None)))
            return
        _capi.check(self._lib.ticit_hip_sample_device(self._handle, shots, seed, first_shot, int(bit_packed),
                                                      int(keep_records), ptr(outputs[0]), ptr(outputs[1]),
                                                      ptr(outputs[2]), ptr(outputs[3]), ptr(ev),
                                                      ctypes.c_void_p(stream or None)))

    def sample_injected(self, shots: int, exo: np.ndarray | None = None, uniforms: np.ndarray | None = None,
                        coins: np.ndarray | None = None, bit_packed: bool = True,
                        want_probs: bool = False) -> SampleResult:
        """Sampling with caller-supplied randomness: exogenous symbol
        bit-planes [nsymbols][ceil(shots/64)], branch uniforms
        [measure_instructions][shots] and coin bits
        [coin_instructions][ceil(shots/64)]."""
        words = (shots + 63) // 64

        def arr(a, shape, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            if a.size != int(np.prod(shape)):
                raise ValueError(f"injected buffer has {a.size} entries, expected {shape}")
            return a

        exo = arr(exo, (self.num_symbols, words), np.uint64)
        uniforms = arr(uniforms, (self.num_measure_instructions, shots), np.float64)
        coins = arr(coins, (self.num_coin_instructions, words), np.uint64)
        probs = np.zeros((self.num_measure_instructions, shots)) if want_probs else None
        res = _capi.Result()
        ptr = lambda a, t: a.ctypes.data_as(t) if a is not None else None
        _capi.check(self._lib.ticit_hip_sample_injected(self._handle, shots, ptr(exo, _capi._u64p),
                                                        ptr(uniforms, _capi._f64p), ptr(coins, _capi._u64p),
                                                        int(bit_packed), ctypes.byref(res),
                                                        ptr(probs, _capi._f64p)))
        try:
            out = _convert(res, True)
        finally:
            self._lib.ticit_hip_result_free(ctypes.byref(res))
        out.probabilities = probs
        return out


    def final_state(self, shots: int = 1, exo=None, uniforms=None, coins=None) -> np.ndarray:
        """Shot 0's final active state (complex128) of an injected run."""
        words = (shots + 63) // 64
        exo = np.ascontiguousarray(exo if exo is not None else np.zeros((max(1, self.num_symbols), words)),
                                   dtype=np.uint64)
        uniforms = np.ascontiguousarray(
            uniforms if uniforms is not None else np.zeros((max(1, self.num_measure_instructions), shots)),
            dtype=np.float64)
        coins = np.ascontiguousarray(
            coins if coins is not None else np.zeros((max(1, self.num_coin_instructions), words)), dtype=np.uint64)
        dim = 1 << self.max_active_qubits
        re = np.zeros(dim)
        im = np.zeros(dim)
        k = ctypes.c_int(0)
        _capi.check(self._lib.ticit_hip_final_state(
            self._handle, shots, exo.ctypes.data_as(_capi._u64p), uniforms.ctypes.data_as(_capi._f64p),
            coins.ctypes.data_as(_capi._u64p), re.ctypes.data_as(_capi._f64p), im.ctypes.data_as(_capi._f64p),
            ctypes.byref(k)))
        d = 1 << k.value
        return re[:d] + 1j * im[:d]


class Circuit:
    """A parsed ``.ticit`` circuit (ticit's Circuit, src/circuit/mod.rs)."""

    def __init__(self, text: str):
        self.text = text
        self.blob = plan_text(text)
        header = self.blob[:10]
        self.num_qubits = int(header[2])
        self.max_active_qubits = int(header[4])
        self.num_symbols = int(header[5])
        self.num_measurements = int(header[6])
        self.num_detectors = int(header[7])

    @classmethod
    def from_file(cls, path: str) -> "Circuit":
        with open(path) as f:
            return cls(f.read())

    def compile(self, postselection_mask=None, *, observable: int = 0, backend: str = "hip",
                precision: str = "fp64", tier: str = "auto", planner: str = "heuristic",
                threads: int | None = None, wide_batch: int = 0,
                components: str = "planned", group_threads: int = 0, pin_measurements=None,
                normalize_syndromes: bool = False, expected_detectors=None,
                expected_observables=None, lane_store: str = "auto") -> CompiledSampler:
        """Circuit::compile with ticit_py's ``backend=`` selector; only the
        ``"hip"`` backend lives in this package."""
        if backend != "hip":
            raise ValueError(f"backend {backend!r} is not provided by ticit_amd; use backend='hip'")
        blob = self.blob
        if postselection_mask is not None and any(int(v) for v in postselection_mask):
            blob = plan_text(self.text, postselection_mask)
        return CompiledSampler(blob, observable=observable, precision=precision, tier=tier,
                               planner=planner, wide_batch=wide_batch, components=components,
                               group_threads=group_threads, pin_measurements=pin_measurements,
                               normalize_syndromes=normalize_syndromes, expected_detectors=expected_detectors,
                               expected_observables=expected_observables, lane_store=lane_store)

    def reference_sample(self) -> ReferenceSample:
        """Circuit::reference_sample (src/circuit/mod.rs:89-91): the
        noiseless detector and observable values, no pins."""
        s = CompiledSampler(self.blob)
        try:
            return s.reference_sample()
        finally:
            s.close()


def sample_batch(samplers, shots, seeds, bit_packed: bool = False, keep_records: bool = True) -> list:
    """ticit_hip_sample_batch: several programs in one batched submission,
    host results (one SampleResult per program)."""
    lib = _capi.load()
    n = len(samplers)
    progs = (ctypes.c_void_p * max(1, n))(*[s.handle.value for s in samplers])
    sh = (ctypes.c_uint64 * max(1, n))(*[int(v) for v in shots])
    sd = (ctypes.c_uint64 * max(1, n))(*[int(v) for v in seeds])
    res = (_capi.Result * max(1, n))()
    _capi.check(lib.ticit_hip_sample_batch(progs, sh, sd, n, int(bit_packed), int(keep_records), res))
    out = []
    try:
        for i in range(n):
            out.append(_convert(res[i], keep_records))
    finally:
        for i in range(n):
            lib.ticit_hip_result_free(ctypes.byref(res[i]))
    return out


class Batch:
    """Many compiled samplers run as one captured hipGraph
    (ticit_hip_batch_*). ``outputs[i]`` holds device buffers (anything with
    ``data_ptr()``, e.g. torch tensors) for item i: measurements,
    detectors, observables ([shots][row] uint8) and observable_ones
    ([nobservables] 64-bit counts, accumulated across launches), then
    optionally exp_vals and, for a postselected sampler, an ``accepted``
    64-bit word that receives its survivor count (survivors' rows first)."""

    def __init__(self, samplers, shots, seeds, outputs, bit_packed: bool = True, keep_records: bool = True,
                 autotune: bool = True, autotune_shots: int = 0, group_grid_percent: int = 0,
                 wide_cu_percent: int = 0):
        lib = _capi.load()
        self._lib = lib
        self._keep = (list(samplers), list(outputs))  # buffers and handles outlive the batch
        n = len(self._keep[0])
        items = (_capi.BatchItem * max(1, n))()
        ptr = lambda t: int(t.data_ptr()) if t is not None else None
        for i, (s, sh, sd, out) in enumerate(zip(samplers, shots, seeds, outputs)):
            items[i].program = s.handle.value
            items[i].shots = int(sh)
            items[i].seed = int(sd)
            items[i].measurements = ptr(out[0])
            items[i].detectors = ptr(out[1])
            items[i].observables = ptr(out[2])
            items[i].observable_ones = ptr(out[3])
            items[i].exp_vals = ptr(out[4]) if len(out) > 4 else None
            items[i].accepted = ptr(out[5]) if len(out) > 5 else None
        opt = _capi.BatchOptions()
        opt.bit_packed = int(bit_packed)
        opt.keep_records = int(keep_records)
        opt.autotune = int(autotune)
        opt.autotune_shots = int(autotune_shots)
        opt.group_grid_percent = int(group_grid_percent)
        opt.wide_cu_percent = int(wide_cu_percent)
        handle = ctypes.c_void_p()
        _capi.check(lib.ticit_hip_batch_create(items, n, ctypes.byref(opt), ctypes.byref(handle)))
        self._handle = handle

    def launch(self, seed_offset: int = 0, stream: int = 0) -> None:
        """Enqueue one replay on ``stream`` (a hipStream_t as an int; 0 =
        default stream); every item draws from seed + seed_offset."""
        _capi.check(self._lib.ticit_hip_batch_launch(self._handle, int(seed_offset), ctypes.c_void_p(stream or None)))

    def wait(self) -> None:
        _capi.check(self._lib.ticit_hip_batch_wait(self._handle))

    def info(self) -> dict:
        inf = _capi.BatchInfo()
        _capi.check(self._lib.ticit_hip_batch_info_get(self._handle, ctypes.byref(inf)))
        nb = min(16, int(inf.group_buckets))
        return dict(items=int(inf.items), lane_items=int(inf.lane_items), wide_items=int(inf.wide_items),
                    buckets=[dict(threads_per_shot=int(inf.bucket_tps[j]), items=int(inf.bucket_items[j]),
                                  region_bytes=int(inf.bucket_region_bytes[j]), grid=int(inf.bucket_grid[j]),
                                  tuned_ms=[float(v) for v in inf.bucket_tuned_ms[j]]) for j in range(nb)])

    def reference_sample(self) -> "ReferenceSample":
        """The noiseless reference sample of this program, obeying its pins
        (reference_sample_for_program, src/sampler/prepared.rs:355-405),
        computed on the device."""
        det = np.zeros(max(1, self.num_detectors), dtype=np.uint8)
        obs = np.zeros(max(1, self.num_observables), dtype=np.uint8)
        _capi.check(self._lib.ticit_hip_reference_sample(self._handle, det.ctypes.data, obs.ctypes.data))
        return ReferenceSample([bool(v) for v in det[:self.num_detectors]],
                               [bool(v) for v in obs[:self.num_observables]])

    def close(self) -> None:
        if getattr(self, "_handle", None):
            self._lib.ticit_hip_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


FEATURE_NAMES = [
    "qubits", "initial_k", "max_k", "symbols", "records", "detectors", "instructions", "measurements",
    "rotations", "promotions", "quantum_ops", "expression_terms", "selected", "planned", "dense_work",
    "component_work", "dense_peak", "peak_live", "allocated", "max_component_k", "merges",
    "dense_amp_traffic", "component_amp_traffic", "output_bytes",
]


def plan_features(blob: np.ndarray) -> dict:
    """Host-side planning features of a plan blob (no device needed): the
    component cost model's quantities (src/component_plan.rs:87-104) and the
    program's shape."""
    lib = _capi.load()
    blob = np.ascontiguousarray(blob, dtype=np.uint64)
    out = np.zeros(len(FEATURE_NAMES))
    _capi.check(lib.ticit_hip_plan_features(blob.ctypes.data_as(_capi._u64p), blob.size,
                                            out.ctypes.data_as(_capi._f64p), out.size))
    return dict(zip(FEATURE_NAMES, out.tolist()))


def device_properties() -> dict:
    props = (ctypes.c_int64 * 6)()
    n = _capi.load().ticit_hip_device_query(props)
    keys = ["lds_per_block", "compute_units", "clock_khz", "warp_size", "memory_clock_khz", "bus_width_bits"]
    return dict(devices=int(n), **{k: int(v) for k, v in zip(keys, props)})
