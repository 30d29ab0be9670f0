"""ctypes binding of ``libticit_hip.so`` (include/ticit_hip.h).

The library is built in-tree by ``ticit_amd/build.py``; loading fails loudly
when it is missing, so no Python path ever stands in for the HIP backend.
"""
from __future__ import annotations

import ctypes
import os

from . import build as _build

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)


class Options(ctypes.Structure):
    _fields_ = [("observable", ctypes.c_uint32), ("precision", ctypes.c_uint32),
                ("planner_mode", ctypes.c_uint32), ("tier", ctypes.c_uint32),
                ("wide_batch", ctypes.c_uint32), ("components", ctypes.c_uint32),
                ("group_threads", ctypes.c_uint32), ("lane_store", ctypes.c_uint32)]


class Result(ctypes.Structure):
    _fields_ = [("shots", ctypes.c_uint64), ("discarded", ctypes.c_uint64),
                ("accepted", ctypes.c_uint64), ("logical_errors", ctypes.c_uint64),
                ("record_rows", ctypes.c_uint64), ("bit_packed", ctypes.c_int32),
                ("reserved0", ctypes.c_int32), ("nmeasurements", ctypes.c_uint64),
                ("ndetectors", ctypes.c_uint64), ("nobservables", ctypes.c_uint64),
                ("measurements", _u8p), ("detectors", _u8p), ("observables", _u8p),
                ("observable_ones", _u64p), ("sample_s", ctypes.c_double),
                ("kernel_s", ctypes.c_double), ("nexpvals", ctypes.c_uint64), ("exp_vals", _f64p)]


class ProgramInfo(ctypes.Structure):
    _fields_ = [("qubits", ctypes.c_uint64), ("measurement_records", ctypes.c_uint64),
                ("detectors", ctypes.c_uint64), ("observables", ctypes.c_uint64),
                ("max_active_qubits", ctypes.c_uint64), ("symbols", ctypes.c_uint64),
                ("instructions", ctypes.c_uint64), ("measure_instructions", ctypes.c_uint64),
                ("coin_instructions", ctypes.c_uint64), ("tier", ctypes.c_uint32),
                ("precision", ctypes.c_uint32), ("wide_batch", ctypes.c_uint32),
                ("components", ctypes.c_uint32), ("output_bytes_per_shot", ctypes.c_double),
                ("state_bytes_per_shot", ctypes.c_double), ("algorithmic_bytes_per_shot", ctypes.c_double),
                ("state_dim", ctypes.c_uint64), ("component_count", ctypes.c_uint64),
                ("max_local_k", ctypes.c_int32), ("planner_source", ctypes.c_int32),
                ("planner_p_components", ctypes.c_double), ("lds_bytes_per_block", ctypes.c_uint32),
                ("blocks_per_cu", ctypes.c_uint32), ("table_words", ctypes.c_uint32 * 4),
                ("expectation_values", ctypes.c_uint64), ("postselected_detectors", ctypes.c_uint64),
                ("lane_store", ctypes.c_uint32), ("lane_region_bytes", ctypes.c_uint32)]


class Parity(ctypes.Structure):
    _fields_ = [("records", ctypes.POINTER(ctypes.c_uint64)), ("nrecords", ctypes.c_size_t),
                ("value", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ReferenceBits(ctypes.Structure):
    _fields_ = [("detectors", ctypes.POINTER(ctypes.c_uint8)), ("ndetectors", ctypes.c_size_t),
                ("observables", ctypes.POINTER(ctypes.c_uint8)), ("nobservables", ctypes.c_size_t),
                ("normalize_syndromes", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class BatchItem(ctypes.Structure):
    _fields_ = [("program", ctypes.c_void_p), ("shots", ctypes.c_uint64), ("seed", ctypes.c_uint64),
                ("measurements", ctypes.c_void_p), ("detectors", ctypes.c_void_p),
                ("observables", ctypes.c_void_p), ("observable_ones", ctypes.c_void_p),
                ("exp_vals", ctypes.c_void_p), ("accepted", ctypes.c_void_p)]


class BatchOptions(ctypes.Structure):
    _fields_ = [("bit_packed", ctypes.c_int32), ("keep_records", ctypes.c_int32), ("autotune", ctypes.c_uint32),
                ("autotune_shots", ctypes.c_uint32), ("group_grid_percent", ctypes.c_uint32),
                ("wide_cu_percent", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 2)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("items", ctypes.c_uint64), ("lane_items", ctypes.c_uint64), ("wide_items", ctypes.c_uint64),
                ("group_buckets", ctypes.c_uint64), ("bucket_tps", ctypes.c_uint32 * 16),
                ("bucket_items", ctypes.c_uint32 * 16), ("bucket_region_bytes", ctypes.c_uint32 * 16),
                ("bucket_grid", ctypes.c_uint32 * 16), ("bucket_tuned_ms", (ctypes.c_double * 8) * 16)]


# ticit_hip_status -> exception kinds of TicitError (src/errors.rs:9-57)
STATUS_NAMES = {1: "InvalidInput", 2: "Parse", 3: "Unsupported", 4: "Internal",
                5: "OutOfDeviceMemory", 6: "DeviceError"}

# every symbol include/ticit_hip.h declares
EXPORTS = [
    "ticit_hip_last_error", "ticit_hip_plan_text", "ticit_hip_plan_text_postselected", "ticit_hip_blob_free",
    "ticit_hip_program_create", "ticit_hip_program_create_pinned", "ticit_hip_program_create_ex",
    "ticit_hip_reference_sample", "ticit_hip_compile_text",
    "ticit_hip_program_destroy",
    "ticit_hip_program_info_get", "ticit_hip_sample", "ticit_hip_sample_device",
    "ticit_hip_sample_device_postselected", "ticit_hip_sample_device_postselected_noalloc",
    "ticit_hip_sample_injected", "ticit_hip_final_state", "ticit_hip_sample_batch",
    "ticit_hip_result_free", "ticit_hip_plan_features",
    "ticit_hip_batch_create", "ticit_hip_batch_launch", "ticit_hip_batch_wait", "ticit_hip_batch_info_get",
    "ticit_hip_batch_destroy", "ticit_hip_planner_features", "ticit_hip_mlp_create", "ticit_hip_mlp_train",
    "ticit_hip_mlp_gradients", "ticit_hip_mlp_predict", "ticit_hip_mlp_params_get", "ticit_hip_mlp_destroy",
    "ticit_hip_planner_set",
    "ticit_hip_device_query",
]

_LIB = None


def library_path() -> str:
    return _build.LIB


def load() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"HIP backend library missing: {path} (run ticit_amd/build.py)")
    # torch ships its own libamdhip64.so.7. Importing it first makes the
    # dynamic linker resolve this library's libamdhip64.so.7 dependency to
    # the runtime already mapped, so the process holds one HIP runtime and
    # torch-allocated device buffers and streams are valid here.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    lib.ticit_hip_last_error.restype = ctypes.c_char_p
    lib.ticit_hip_plan_text_postselected.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p,
                                                    ctypes.c_size_t, ctypes.POINTER(_u64p),
                                                    ctypes.POINTER(ctypes.c_size_t)]
    lib.ticit_hip_program_create_pinned.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_void_p,
                                                    ctypes.POINTER(Parity), ctypes.c_size_t,
                                                    ctypes.POINTER(ctypes.c_void_p)]
    lib.ticit_hip_program_create_ex.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_void_p,
                                                ctypes.POINTER(Parity), ctypes.c_size_t,
                                                ctypes.POINTER(ReferenceBits), ctypes.POINTER(ctypes.c_void_p)]
    lib.ticit_hip_reference_sample.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ticit_hip_plan_text.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_u64p),
                                        ctypes.POINTER(ctypes.c_size_t)]
    lib.ticit_hip_blob_free.argtypes = [_u64p]
    lib.ticit_hip_program_create.argtypes = [_u64p, ctypes.c_size_t, ctypes.POINTER(Options),
                                             ctypes.POINTER(ctypes.c_void_p)]
    lib.ticit_hip_compile_text.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(Options),
                                           ctypes.POINTER(ctypes.c_void_p)]
    lib.ticit_hip_program_destroy.argtypes = [ctypes.c_void_p]
    lib.ticit_hip_program_info_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(ProgramInfo)]
    lib.ticit_hip_sample.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                     ctypes.c_int, ctypes.POINTER(Result)]
    lib.ticit_hip_sample_device.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                            ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]
    lib.ticit_hip_sample_device_postselected.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                                         ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                         ctypes.c_void_p]
    lib.ticit_hip_sample_device_postselected_noalloc.argtypes = \
        lib.ticit_hip_sample_device_postselected.argtypes
    lib.ticit_hip_sample_injected.argtypes = [ctypes.c_void_p, ctypes.c_uint64, _u64p, _f64p, _u64p,
                                              ctypes.c_int, ctypes.POINTER(Result), _f64p]
    lib.ticit_hip_final_state.argtypes = [ctypes.c_void_p, ctypes.c_uint64, _u64p, _f64p, _u64p, _f64p, _f64p,
                                          ctypes.POINTER(ctypes.c_int)]
    lib.ticit_hip_sample_batch.argtypes = [ctypes.POINTER(ctypes.c_void_p), _u64p, _u64p, ctypes.c_size_t,
                                           ctypes.c_int, ctypes.c_int, ctypes.POINTER(Result)]
    lib.ticit_hip_result_free.argtypes = [ctypes.POINTER(Result)]
    lib.ticit_hip_plan_features.argtypes = [_u64p, ctypes.c_size_t, _f64p, ctypes.c_size_t]
    lib.ticit_hip_batch_create.argtypes = [ctypes.POINTER(BatchItem), ctypes.c_size_t, ctypes.POINTER(BatchOptions),
                                           ctypes.POINTER(ctypes.c_void_p)]
    lib.ticit_hip_batch_launch.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    lib.ticit_hip_batch_wait.argtypes = [ctypes.c_void_p]
    lib.ticit_hip_batch_info_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchInfo)]
    lib.ticit_hip_batch_destroy.argtypes = [ctypes.c_void_p]
    lib.ticit_hip_batch_destroy.restype = None
    _f32p = ctypes.POINTER(ctypes.c_float)
    lib.ticit_hip_planner_features.argtypes = [_u64p, ctypes.c_size_t, _f32p]
    lib.ticit_hip_mlp_create.argtypes = [_f32p, ctypes.POINTER(ctypes.c_void_p)]
    lib.ticit_hip_mlp_train.argtypes = [ctypes.c_void_p, _f32p, ctypes.POINTER(ctypes.c_int32), _f32p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_uint32), ctypes.c_size_t, ctypes.c_uint32,
                                        ctypes.c_float, ctypes.c_float, _f32p]
    lib.ticit_hip_mlp_gradients.argtypes = [ctypes.c_void_p, _f32p, ctypes.POINTER(ctypes.c_int32), _f32p,
                                            ctypes.c_size_t, ctypes.c_float, _f32p, _f32p]
    lib.ticit_hip_mlp_predict.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_size_t, _f32p]
    lib.ticit_hip_mlp_params_get.argtypes = [ctypes.c_void_p, _f32p]
    lib.ticit_hip_mlp_destroy.argtypes = [ctypes.c_void_p]
    lib.ticit_hip_mlp_destroy.restype = None
    lib.ticit_hip_planner_set.argtypes = [_f32p, _f32p, _f32p, ctypes.c_float]
    lib.ticit_hip_device_query.argtypes = [ctypes.POINTER(ctypes.c_int64)]
    lib.ticit_hip_device_query.restype = ctypes.c_int
    for name in ("ticit_hip_plan_text", "ticit_hip_program_create", "ticit_hip_compile_text",
                 "ticit_hip_program_info_get", "ticit_hip_sample", "ticit_hip_sample_device",
                 "ticit_hip_sample_device_postselected", "ticit_hip_sample_device_postselected_noalloc",
                 "ticit_hip_sample_injected", "ticit_hip_sample_batch", "ticit_hip_final_state",
                 "ticit_hip_plan_features", "ticit_hip_batch_create", "ticit_hip_batch_launch",
                 "ticit_hip_batch_wait", "ticit_hip_batch_info_get", "ticit_hip_planner_features",
                 "ticit_hip_mlp_create", "ticit_hip_mlp_train", "ticit_hip_mlp_gradients", "ticit_hip_mlp_predict",
                 "ticit_hip_mlp_params_get", "ticit_hip_planner_set"):
        getattr(lib, name).restype = ctypes.c_int
    lib.ticit_hip_blob_free.restype = None
    lib.ticit_hip_program_destroy.restype = None
    lib.ticit_hip_result_free.restype = None
    _LIB = lib
    return lib


class TicitError(RuntimeError):
    """Raised for a failing status; ``kind`` names the TicitError variant."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status
        self.kind = STATUS_NAMES.get(status, "Internal")


class InvalidInputError(TicitError, ValueError):
    """InvalidInput and Unsupported statuses: a ValueError, as ticit_py
    raises for them (ticit_py/src/lib.rs:24-35)."""


class ParseError(TicitError, ValueError):
    """A circuit failed to parse (ticit_py/src/lib.rs:17-22)."""


_ERROR_CLASSES = {1: InvalidInputError, 2: ParseError, 3: InvalidInputError}


def check(status: int) -> None:
    if status != 0:
        cls = _ERROR_CLASSES.get(status, TicitError)
        raise cls(status, load().ticit_hip_last_error().decode(errors="replace"))
