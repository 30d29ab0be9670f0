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


_PAULI = {"X": GATES1["X"], "Y": GATES1["Y"], "Z": GATES1["Z"]}


def _apply_product(psi: np.ndarray, n: int, product: str) -> tuple[np.ndarray, float]:
    """P|psi> for a product such as ``X0*!Z1``; returns (P psi, sign) where
    each ``!`` flips the sign."""
    sign = 1.0
    out = psi
    for f in product.split("*"):
        while f.startswith("!"):
            sign = -sign
            f = f[1:]
        out = apply1(out, n, int(f[1:]), _PAULI[f[0].upper()])
    return out, sign


def expectation_values(text: str) -> np.ndarray:
    """Exact <psi|P|psi> of every EXP_VAL product, in slot order, for a
    noiseless circuit whose probes all come before its measurements."""
    lines = [l.split("#")[0].strip() for l in text.splitlines()]
    lines = [l for l in lines if l and l != "TICK"]
    n = 1
    for l in lines:
        for tok in l.split()[1:]:
            for f in tok.split("*"):
                f = f.lstrip("!")
                if f[:1].isalpha():
                    f = f[1:]
                if f.isdigit():
                    n = max(n, int(f) + 1)
    psi = np.zeros(2 ** n, dtype=complex)
    psi[0] = 1.0
    vals = []
    measured = False
    for l in lines:
        parts = l.split()
        op = parts[0]
        if op == "EXP_VAL":
            if measured:
                raise ValueError("probes after a measurement depend on the branch")
            for prod in parts[1:]:
                q, sign = _apply_product(psi, n, prod)
                vals.append(sign * float(np.vdot(psi, q).real))
            continue
        if op == "M":
            measured = True
            continue
        targets = [int(t) for t in parts[1:]]
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
    return np.array(vals)


def with_probes(text: str, rng: np.random.Generator, nprobes: int) -> str:
    """Inserts EXP_VAL lines with random Pauli products (random inversions)
    before random gate lines of a circuit whose last line is its only M."""
    lines = [l for l in text.splitlines() if l.strip()]
    n = 1 + max(int(t) for l in lines for t in l.split()[1:] if t.isdigit())
    body, last = lines[:-1], lines[-1]
    for _ in range(nprobes):
        at = int(rng.integers(0, len(body) + 1))
        prods = []
        for _ in range(int(rng.integers(1, 4))):
            qs = rng.choice(n, size=int(rng.integers(1, min(n, 3) + 1)), replace=False)
            f = [("!" if rng.random() < 0.2 else "") + "XYZ"[int(rng.integers(3))] + str(int(q)) for q in qs]
            prods.append("*".join(f))
        body.insert(at, "EXP_VAL " + " ".join(prods))
    return "\n".join(body + [last]) + "\n"
