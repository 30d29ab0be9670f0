"""Plan-blob builder for tests: encodes hand-written instruction streams in
the u64 format of include/ticit_plan.h (version 2), so tests can restate the
reference's hand-built FactoredInstructionPrograms (e.g. the component
planner tests of src/component_plan.rs:560-734)."""
from __future__ import annotations

import struct

import numpy as np

MAGIC = 0x4E4C505449434954
VERSION = 2
ROTATE, PROMOTE, RECORD, DETECTOR, MEASURE, COIN = range(6)


def _f64(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def rotation(nq: int, xmask: int = 0, zmask: int = 0, angle: float = 0.0, phase: int = 0,
             conds=(), constant: bool = False) -> dict:
    return dict(op=ROTATE, nq=nq, xmask=xmask, zmask=zmask, angle=angle, phase=phase,
                conds=list(conds), constant=constant)


def promotion(nq: int, angle: float, conds=(), constant: bool = False) -> dict:
    return dict(op=PROMOTE, nq=nq, angle=angle, conds=list(conds), constant=constant)


def measurement(nq: int, xmask: int = 0, zmask: int = 0, branch: int = 0, record: int = 0,
                phase: int = 0, conds=None) -> dict:
    """PrecomputedActivePauliMeasurementKernel::from_pauli pivot choice
    (src/active.rs:273-285): highest X bit, else highest Z bit; outcome is
    the branch symbol unless ``conds`` says otherwise."""
    if xmask:
        pivot, diagonal = xmask.bit_length() - 1, False
    else:
        pivot, diagonal = zmask.bit_length() - 1, True
    return dict(op=MEASURE, nq=nq, xmask=xmask, zmask=zmask, pivot=pivot, diagonal=diagonal,
                diag_phase_bit=1 if (diagonal and phase == 2) else 0, phase=phase, branch=branch,
                record=record, conds=[branch] if conds is None else list(conds))


def detector(det: int, records) -> dict:
    return dict(op=DETECTOR, detector=det, det_records=list(records))


def encode(n: int, initial_k: int, max_k: int, nsymbols: int, nrecords: int, ndetectors: int, instrs,
           nexpvals: int = 0, use_components: bool = False, observables=()) -> np.ndarray:
    w = [MAGIC, VERSION, n, initial_k, max_k, nsymbols, nrecords, ndetectors, nexpvals,
         1 if use_components else 0, len(instrs)]
    for d in instrs:
        w0 = (d["op"] | (int(d.get("diagonal", False)) << 8) | ((d.get("diag_phase_bit", 0) & 1) << 9)
              | (int(d.get("constant", False)) << 10) | ((d.get("nq", 0) & 0xFFFF) << 16)
              | ((d.get("pivot", 0) & 0xFFFF) << 32) | ((d.get("phase", 0) & 3) << 48))
        w += [w0, d.get("xmask", 0), d.get("zmask", 0), _f64(d.get("angle", 0.0)),
              (d.get("branch", 0) & 0xFFFFFFFF) | ((d.get("record", 0) & 0xFFFFFFFF) << 32),
              (d.get("record_condition", 0) & 0xFFFFFFFF) | ((d.get("detector", 0) & 0xFFFFFFFF) << 32),
              d.get("exp_val", -1) & 0xFFFFFFFF]
        conds = d.get("conds", [])
        w += [len(conds), *conds]
        recs = d.get("det_records", [])
        w += [len(recs), *recs]
    w += [0, 0, 0, 0]  # categorical, rare, Bernoulli, low-probability groups
    w.append(len(observables))
    for ob in observables:
        w += [len(ob), *ob]
    w.append(MAGIC)
    return np.array(w, dtype=np.uint64)


def set_component_flag(blob: np.ndarray, on: bool) -> np.ndarray:
    """Copy of a version-2 blob with use_active_components set or cleared."""
    out = np.array(blob, dtype=np.uint64, copy=True)
    if int(out[1]) != VERSION:
        raise ValueError("component flag needs a version-2 blob")
    out[9] = np.uint64((int(out[9]) & ~1) | (1 if on else 0))
    return out
