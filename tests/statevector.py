"""Exact statevector probabilities for small noiseless circuits (test-only).

Used to pin the front end and both samplers against the mathematics: the
distribution of the final measurement record of a circuit whose only
measurements are the last ``M`` line is |<b|U|0>|^2 over basis states b.
"""
from __future__ import annotations

import numpy as np

_H = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
_S = np.diag([1, 1j])
_T = np.diag([1, np.exp(1j * np.pi / 4)])
GATES1 = {
    "H": _H, "S": _S, "S_DAG": _S.conj(), "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.diag([1, -1]).astype(complex),
    "T": _T, "T_DAG": _T.conj(),
}


def apply1(psi: np.ndarray, n: int, q: int, u: np.ndarray) -> np.ndarray:
    psi = psi.reshape([2] * n)
    ax = n - 1 - q  # qubit q is bit q of the basis index
    psi = np.moveaxis(np.tensordot(u, psi, axes=([1], [ax])), 0, ax)
    return psi.reshape(-1)


def apply_cx(psi: np.ndarray, n: int, c: int, t: int) -> np.ndarray:
    idx = np.arange(psi.size)
    ctrl = (idx >> c) & 1
    out = psi.copy()
    src = idx ^ (ctrl << t)
    out[idx] = psi[src]
    return out


def apply_cz(psi: np.ndarray, n: int, a: int, b: int) -> np.ndarray:
    idx = np.arange(psi.size)
    sign = 1 - 2 * (((idx >> a) & 1) & ((idx >> b) & 1))
    return psi * sign


def final_distribution(text: str) -> tuple[list[int], np.ndarray]:
    """(measured qubits in record order, probability of each record value)
    for a circuit whose last non-annotation line is its only M."""
    lines = [l.split("#")[0].strip() for l in text.splitlines()]
    lines = [l for l in lines if l and l != "TICK"]
    qubits = set()
    for l in lines:
        for tok in l.split()[1:]:
            if tok.isdigit():
                qubits.add(int(tok))
    n = max(qubits) + 1
    psi = np.zeros(2 ** n, dtype=complex)
    psi[0] = 1.0
    measured = None
    for l in lines:
        parts = l.split()
        op, targets = parts[0], [int(t) for t in parts[1:]]
        if op == "M":
            if measured is not None:
                raise ValueError("only one final M line is supported")
            measured = targets
            continue
        if measured is not None:
            raise ValueError("gates after the final measurement")
        if op in GATES1:
            for q in targets:
                psi = apply1(psi, n, q, GATES1[op])
        elif op in ("CX", "CNOT"):
            for i in range(0, len(targets), 2):
                psi = apply_cx(psi, n, targets[i], targets[i + 1])
        elif op == "CZ":
            for i in range(0, len(targets), 2):
                psi = apply_cz(psi, n, targets[i], targets[i + 1])
        else:
            raise ValueError(f"unsupported gate {op}")
    probs = np.abs(psi) ** 2
    m = len(measured)
    out = np.zeros(2 ** m)
    idx = np.arange(psi.size)
    rec = np.zeros(psi.size, dtype=np.int64)
    for j, q in enumerate(measured):
        rec |= ((idx >> q) & 1) << j
    np.add.at(out, rec, probs)
    return measured, out


def record_values(rows: np.ndarray, nrec: int, packed: bool = True) -> np.ndarray:
    """Integer value of each row's first ``nrec`` record bits."""
    bits = np.unpackbits(rows, axis=1, bitorder="little")[:, :nrec] if packed else rows[:, :nrec]
    return (bits.astype(np.int64) << np.arange(nrec)).sum(axis=1)


def tvd(counts: np.ndarray, probs: np.ndarray) -> float:
    return 0.5 * float(np.abs(counts / counts.sum() - probs).sum())
