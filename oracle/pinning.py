"""Pinned measurement parities, restated for the CPU oracle (test
infrastructure only): the GF(2) solver of ticit's plan_pinned_measurements
(src/sampler/pinning.rs:80-289), written from the reference's description
independently of the product's C++ solver, over the instruction stream of a
plan blob (include/ticit_plan.h).

A record's noiseless value is its outcome expression with every record
condition expanded through the instruction that assigns it, branch symbols
kept and exogenous (noise) symbols dropped. Each constraint row is reduced
against earlier pivot rows, taking as pivot the branch drawn last; the
pivot's instruction then takes the rest of its row instead of a draw.
"""
from __future__ import annotations

import numpy as np

OP_RECORD, OP_DETECTOR, OP_MEASURE, OP_COIN = 2, 3, 4, 5


def decode_instructions(blob: np.ndarray) -> list[dict]:
    """The instruction stream of a plan blob: op, branch, record,
    record_condition, exp_val and the outcome expression."""
    b = [int(x) for x in blob]
    i = 11 if b[1] == 2 else 10
    n = b[i - 1]
    out = []
    for _ in range(n):
        w0 = b[i]
        ev = b[i + 6] & 0xFFFFFFFF
        ncond = b[i + 7]
        conds = b[i + 8:i + 8 + ncond]
        out.append(dict(op=w0 & 0xFF, constant=bool((w0 >> 10) & 1), branch=b[i + 4] & 0xFFFFFFFF,
                        record=b[i + 4] >> 32, rc=b[i + 5] & 0xFFFFFFFF,
                        exp_val=ev if ev < 0x80000000 else ev - (1 << 32), conds=conds))
        i += 8 + ncond
        i += 1 + b[i]
    return out


class PinError(ValueError):
    pass


def plan_pinned(blob: np.ndarray, constraints) -> list[tuple[int, bool, list[int]]]:
    """(instruction, constant, branch symbols) per pinned branch, ordered by
    instruction; constraints are (zero-based records, value) pairs."""
    if not constraints:
        return []
    ins = decode_instructions(blob)
    branch_at, derived_at, record_at = {}, {}, {}
    for idx, x in enumerate(ins):
        if x["exp_val"] >= 0:
            continue
        if x["op"] in (OP_MEASURE, OP_COIN) and x["branch"]:
            branch_at[x["branch"]] = idx
        if x["rc"]:
            derived_at.setdefault(x["rc"], idx)
        if x["record"]:
            record_at[x["record"]] = idx
    memo: dict[int, frozenset] = {}
    constant_of: dict[int, bool] = {}

    def expand(sym: int) -> tuple[bool, frozenset]:
        # noiseless value of a symbol: (constant, set of branch symbols)
        if sym in memo:
            return constant_of[sym], memo[sym]
        if sym in branch_at:
            res = (False, frozenset([sym]))
        elif sym in derived_at:
            x = ins[derived_at[sym]]
            c, s = x["constant"], frozenset()
            for t in x["conds"]:
                tc, ts = expand(t)
                c ^= tc
                s = s ^ ts
            res = (c, s)
        else:
            res = (False, frozenset())  # exogenous noise: zero without noise
        constant_of[sym], memo[sym] = res
        return res

    pivots: dict[int, tuple[int, bool, frozenset]] = {}  # instruction -> (pivot, constant, row)
    for records, value in constraints:
        c, row = bool(value), frozenset()
        for r in records:
            if (r + 1) not in record_at:
                raise PinError(f"pinned measurement record {r} is not written by this circuit")
            x = ins[record_at[r + 1]]
            c ^= x["constant"]
            for t in x["conds"]:
                tc, ts = expand(t)
                c ^= tc
                row = row ^ ts
        while True:
            drawn = [(branch_at[s], s) for s in row if s in branch_at]
            if not drawn:
                if c:
                    raise PinError(f"pinned measurement parity over {list(records)} is deterministic "
                                   f"and cannot be {int(bool(value))}")
                break
            instr, pivot = max(drawn)
            if instr in pivots:
                _, pc, prow = pivots[instr]
                c ^= pc
                row = row ^ prow
            else:
                pivots[instr] = (pivot, c, row)
                break
    return [(instr, c, sorted(row - {pivot})) for instr, (pivot, c, row) in sorted(pivots.items())]
