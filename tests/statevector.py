"""Exact statevector probabilities for small circuits (test-only).

Used to pin the front end and both samplers against the mathematics: the
distribution of the final measurement record of a circuit whose only
measurements are the last ``M`` line is |<b|U|0>|^2 over basis states b;
``record_distribution`` extends this to mid-circuit measurements, resets and
a few Pauli-noise sites as an exact mixture over noise and branch outcomes.
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


# ---------------------------------------------------------------------------
# Noisy circuits with mid-circuit measurements: an exact mixture
# ---------------------------------------------------------------------------

_NOISE1 = {"X_ERROR": "X", "Y_ERROR": "Y", "Z_ERROR": "Z"}
_ONE_Q_PAULIS = [None, "X", "Y", "Z"]


def _pauli_terms(op: str, p: float, targets: list[int]) -> list[list[tuple[float, list[tuple[int, str]]]]]:
    """Per noise site of one line: the (probability, [(qubit, Pauli)]) terms
    of its channel (p/3 per Pauli for DEPOLARIZE1, p/15 per non-identity pair
    for DEPOLARIZE2)."""
    sites = []
    if op in _NOISE1:
        for q in targets:
            sites.append([(1.0 - p, []), (p, [(q, _NOISE1[op])])])
    elif op == "DEPOLARIZE1":
        for q in targets:
            sites.append([(1.0 - p, [])] + [(p / 3.0, [(q, a)]) for a in "XYZ"])
    elif op == "DEPOLARIZE2":
        for i in range(0, len(targets), 2):
            a, b = targets[i], targets[i + 1]
            terms = [(1.0 - p, [])]
            for pa in _ONE_Q_PAULIS:
                for pb in _ONE_Q_PAULIS:
                    if pa is None and pb is None:
                        continue
                    ops = ([(a, pa)] if pa else []) + ([(b, pb)] if pb else [])
                    terms.append((p / 15.0, ops))
            sites.append(terms)
    else:
        raise ValueError(op)
    return sites


def record_distribution(text: str, max_noise_sites: int = 3) -> np.ndarray:
    """Exact probability of every measurement-record value (bit j = record j)
    of a small circuit with mid-circuit M, R, MR, Pauli-error channels
    (X/Y/Z_ERROR, DEPOLARIZE1, DEPOLARIZE2) and unitary gates, as a mixture:
    every noise site splits each branch by its Pauli terms, every
    measurement by its Born probabilities. Detector and observable lines are
    ignored. At most `max_noise_sites` noise sites (one per target, or per
    target pair for DEPOLARIZE2)."""
    lines = [l.split("#")[0].strip() for l in text.splitlines()]
    lines = [l for l in lines if l and l != "TICK"]
    n = 1
    for l in lines:
        head = l.split()[0]
        if head.startswith(("DETECTOR", "OBSERVABLE_INCLUDE")):
            continue
        for tok in l.split()[1:]:
            if tok.isdigit():
                n = max(n, int(tok) + 1)
    psi0 = np.zeros(2 ** n, dtype=complex)
    psi0[0] = 1.0
    branches = [(1.0, psi0, 0, 0)]  # (weight, normalized state, record bits, record count)
    nsites = 0
    idx = np.arange(2 ** n)

    def measure(brs, q, reset):
        out = []
        for w, psi, rec, nr in brs:
            one = ((idx >> q) & 1) == 1
            p1 = float(np.sum(np.abs(psi[one]) ** 2))
            for bit, pb in ((0, 1.0 - p1), (1, p1)):
                if pb <= 1e-15:
                    continue
                phi = np.where(one if bit else ~one, psi, 0) / np.sqrt(pb)
                if reset and bit:
                    phi = apply1(phi, n, q, GATES1["X"])
                out.append((w * pb, phi, rec | (bit << nr), nr + 1))
        return out

    for l in lines:
        parts = l.split()
        head = parts[0]
        if head.startswith(("DETECTOR", "OBSERVABLE_INCLUDE")):
            continue
        op, _, arg = head.partition("(")
        targets = [int(t) for t in parts[1:]]
        if op in ("M", "MR"):
            for q in targets:
                branches = measure(branches, q, op == "MR")
        elif op == "R":
            for q in targets:
                # reset = measure without a record, then flip a 1 back to 0
                out = []
                for w, psi, rec, nr in branches:
                    for w2, phi, _, _ in measure([(w, psi, 0, 0)], q, True):
                        out.append((w2, phi, rec, nr))
                branches = out
        elif op in _NOISE1 or op in ("DEPOLARIZE1", "DEPOLARIZE2"):
            p = float(arg.rstrip(")"))
            for terms in _pauli_terms(op, p, targets):
                nsites += 1
                if nsites > max_noise_sites:
                    raise ValueError("too many noise sites for an exact mixture")
                out = []
                for w, psi, rec, nr in branches:
                    for pt, ops in terms:
                        if pt <= 0.0:
                            continue
                        phi = psi
                        for q, a in ops:
                            phi = apply1(phi, n, q, _PAULI[a])
                        out.append((w * pt, phi, rec, nr))
                branches = out
        elif op in GATES1:
            branches = [(w, psi_apply(psi, n, op, targets), rec, nr) for w, psi, rec, nr in branches]
        elif op in ("CX", "CNOT", "CZ"):
            branches = [(w, psi_apply(psi, n, op, targets), rec, nr) for w, psi, rec, nr in branches]
        else:
            raise ValueError(f"unsupported instruction {op}")
    nrec = branches[0][3] if branches else 0
    dist = np.zeros(2 ** nrec)
    for w, _, rec, nr in branches:
        assert nr == nrec
        dist[rec] += w
    return dist


def psi_apply(psi: np.ndarray, n: int, op: str, targets: list[int]) -> np.ndarray:
    if op in GATES1:
        for q in targets:
            psi = apply1(psi, n, q, GATES1[op])
        return psi
    for i in range(0, len(targets), 2):
        psi = (apply_cz if op == "CZ" else apply_cx)(psi, n, targets[i], targets[i + 1])
    return psi


def chi2_pvalue(counts: np.ndarray, probs: np.ndarray, min_expected: float = 5.0) -> float:
    """Pearson chi-square goodness of fit of `counts` to `probs`; outcomes
    whose expected count is below `min_expected` are pooled into one bin."""
    from scipy.stats import chi2
    counts = np.asarray(counts, dtype=float)
    exp = np.asarray(probs, dtype=float) * counts.sum()
    big = exp >= min_expected
    obs_b, exp_b = list(counts[big]), list(exp[big])
    if (~big).any() and exp[~big].sum() > 0:
        obs_b.append(counts[~big].sum())
        exp_b.append(exp[~big].sum())
    obs_b, exp_b = np.array(obs_b), np.array(exp_b)
    stat = float(((obs_b - exp_b) ** 2 / exp_b).sum())
    dof = max(1, len(obs_b) - 1)
    return float(chi2.sf(stat, dof))


def noisy_midcircuit_circuits(count: int = 6, seed: int = 31) -> list[str]:
    """Small random circuits: one-qubit Cliffords and T, CX layers, a
    mid-circuit M (or MR) on some qubits, three noise sites drawn from
    X_ERROR, Y_ERROR, DEPOLARIZE1 and DEPOLARIZE2 at probabilities that
    visibly move the distribution, and a final M on every qubit."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        n = 4 + c % 2
        sites = 0
        lines = []
        for layer in range(6):
            for q in range(n):
                g = rng.choice(["H", "S", "T", "T_DAG", "H", ""])
                if g:
                    lines.append(f"{g} {q}")
            perm = rng.permutation(n)
            lines.append("CX " + " ".join(f"{perm[2 * i]} {perm[2 * i + 1]}" for i in range(n // 2)))
            if layer in (1, 3, 4) and sites < 3:
                kind = ["X_ERROR(0.2)", "Y_ERROR(0.15)", "DEPOLARIZE1(0.25)", "DEPOLARIZE2(0.3)"][(c + layer) % 4]
                qs = rng.choice(n, size=2, replace=False)
                lines.append(f"{kind} {qs[0]} {qs[1]}" if kind.startswith("DEPOLARIZE2") else f"{kind} {qs[0]}")
                sites += 1
            if layer == 2:
                lines.append(("MR " if c % 3 == 2 else "M ") + " ".join(str(q) for q in range(0, n, 2)))
        lines.append("M " + " ".join(str(q) for q in range(n)))
        out.append("\n".join(lines) + "\n")
    return out
