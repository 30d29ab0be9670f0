"""Exact statevector probabilities for small circuits (test-only).

Used to pin the front end and both samplers against the mathematics: the
distribution of the final measurement record of a circuit whose only
measurements are the last ``M`` line is |<b|U|0>|^2 over basis states b;
``record_distribution`` extends this to mid-circuit measurements, resets and
a few Pauli-noise sites as an exact mixture over noise and branch outcomes.
"""
from __future__ import annotations

import math

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


def _product(tokens, n):
    """[(qubit, axis)] and sign of a Pauli product written as tokens such as
    ``X0*!Z1`` (one token) or ``X0 Z1`` (one factor per token)."""
    ops, sign = [], 1.0
    for tok in tokens:
        for f in tok.split("*"):
            while f.startswith("!"):
                sign = -sign
                f = f[1:]
            ops.append((int(f[1:]), f[0].upper()))
    return ops, sign


def _apply_ops(psi, n, ops):
    for q, a in ops:
        psi = apply1(psi, n, q, _PAULI[a])
    return psi


def _pauli_channel_terms(k: int, probs: list, qubits: list) -> list:
    """PAULI_CHANNEL_k per target group: probabilities in the order of the
    axis codes (I, X, Y, Z per target), the first target's code the most
    significant."""
    axes = [None, "X", "Y", "Z"]
    sites = []
    for i in range(0, len(qubits), k):
        grp = qubits[i:i + k]
        terms = [(1.0 - sum(probs), [])]
        for code in range(1, 4 ** k):
            ops, rem = [], code
            for pos in reversed(range(k)):
                if axes[rem & 3]:
                    ops.append((grp[pos], axes[rem & 3]))
                rem >>= 2
            terms.append((probs[code - 1], ops))
        sites.append(terms)
    return sites


def record_distribution(text: str, max_noise_sites: int = 3) -> np.ndarray:
    """Exact probability of every measurement-record value (bit j = record j)
    of a small circuit, as a mixture: every noise site splits each branch by
    its Pauli terms, every measurement by its Born probabilities.

    Supported: one-qubit gates and CX / CZ; M, MR, R; Pauli-product
    measurements MPP, MXX, MYY, MZZ; Pauli rotations exp(-i a P) for SPP
    (a = pi/4), SPP_DAG (-pi/4) and R_PAULI (half turns * pi/2); MPAD;
    X/Y/Z_ERROR, DEPOLARIZE1 (p/3 per Pauli), DEPOLARIZE2 (p/15 per pair),
    PAULI_CHANNEL_1/2, E / ELSE_CORRELATED_ERROR chains (alternative i fires
    with p_i times the product of (1 - p_j) over earlier ones), and
    HERALDED_ERASE / HERALDED_PAULI_CHANNEL_1 (a herald record, then I, X,
    Y or Z). Detector and observable lines are ignored. At most
    `max_noise_sites` noise sites (one per target, target group or chain)."""
    lines = [l.split("#")[0].strip() for l in text.splitlines()]
    lines = [l for l in lines if l and l != "TICK"]
    n = 1
    for l in lines:
        head = l.split()[0]
        if head.startswith(("DETECTOR", "OBSERVABLE_INCLUDE", "MPAD")):
            continue
        for tok in l.split()[1:]:
            for f in tok.split("*"):
                f = f.lstrip("!")
                if f[:1].upper() in ("X", "Y", "Z"):
                    f = f[1:]
                if f.isdigit():
                    n = max(n, int(f) + 1)
    psi0 = np.zeros(2 ** n, dtype=complex)
    psi0[0] = 1.0
    state = {"branches": [(1.0, psi0, 0, 0)], "sites": 0}  # (weight, normalized state, records, count)
    idx = np.arange(2 ** n)

    def site():
        state["sites"] += 1
        if state["sites"] > max_noise_sites:
            raise ValueError("too many noise sites for an exact mixture")

    def noise(sites_terms):
        for terms in sites_terms:
            site()
            state["branches"] = [(w * pt, _apply_ops(psi, n, ops), rec, nr)
                                 for w, psi, rec, nr in state["branches"] for pt, ops in terms if pt > 0.0]

    def measure_z(q, reset, record=True):
        out = []
        for w, psi, rec, nr in state["branches"]:
            one = ((idx >> q) & 1) == 1
            p1 = float(np.sum(np.abs(psi[one]) ** 2))
            for bit, pb in ((0, 1.0 - p1), (1, p1)):
                if pb <= 1e-15:
                    continue
                phi = np.where(one if bit else ~one, psi, 0) / np.sqrt(pb)
                if reset and bit:
                    phi = apply1(phi, n, q, GATES1["X"])
                out.append((w * pb, phi, rec | (bit << nr) if record else rec, nr + 1 if record else nr))
        state["branches"] = out

    def measure_product(ops, sign):
        out = []
        for w, psi, rec, nr in state["branches"]:
            ppsi = sign * _apply_ops(psi, n, ops)
            for bit in (0, 1):
                proj = 0.5 * (psi + (ppsi if bit == 0 else -ppsi))
                pb = float(np.vdot(proj, proj).real)
                if pb <= 1e-15:
                    continue
                out.append((w * pb, proj / np.sqrt(pb), rec | (bit << nr), nr + 1))
        state["branches"] = out

    chain = []  # open correlated-error alternatives: (p, ops)

    def flush_chain():
        if chain:
            terms, rem = [], 1.0
            for p, ops in chain:
                terms.append((rem * p, ops))
                rem *= 1.0 - p
            terms.append((rem, []))
            chain.clear()
            noise([terms])

    for l in lines:
        parts = l.split()
        head = parts[0]
        op, _, arg = head.partition("(")
        args = [float(v) for v in arg.rstrip(")").split(",")] if arg else []
        if op in ("E", "CORRELATED_ERROR", "ELSE_CORRELATED_ERROR"):
            if op != "ELSE_CORRELATED_ERROR":
                flush_chain()
            chain.append((args[0], _product(parts[1:], n)[0]))
            continue
        flush_chain()
        if head.startswith(("DETECTOR", "OBSERVABLE_INCLUDE")):
            continue
        if op in ("MPP", "MXX", "MYY", "MZZ"):
            if op == "MPP":
                prods = [_product([t], n) for t in parts[1:]]
            else:
                prods = []
                for a, b in zip(parts[1::2], parts[2::2]):
                    ops, _ = _product([op[1] + a.lstrip("!"), op[1] + b.lstrip("!")], n)
                    prods.append((ops, -1.0 if a.startswith("!") != b.startswith("!") else 1.0))
            for ops, sign in prods:
                measure_product(ops, sign)
        elif op in ("SPP", "SPP_DAG", "R_PAULI"):
            ang = math.pi / 4 if op == "SPP" else -math.pi / 4 if op == "SPP_DAG" else args[0] * math.pi / 2
            for t in parts[1:]:
                ops, sign = _product([t], n)
                state["branches"] = [(w, math.cos(ang) * psi - 1j * math.sin(ang) * sign * _apply_ops(psi, n, ops),
                                      rec, nr) for w, psi, rec, nr in state["branches"]]
        elif op == "MPAD":
            # each literal (or its inversion) a record, flipped with the
            # optional probability
            pf = args[0] if args else 0.0
            for t in parts[1:]:
                bit = int(t.lstrip("!")) ^ int(t.startswith("!"))
                if pf > 0.0:
                    site()
                state["branches"] = [(w * pb, psi, rec | ((bit ^ f) << nr), nr + 1)
                                     for w, psi, rec, nr in state["branches"]
                                     for f, pb in ((0, 1.0 - pf), (1, pf)) if pb > 0.0]
        elif op in ("PAULI_CHANNEL_1", "PAULI_CHANNEL_2"):
            noise(_pauli_channel_terms(1 if op == "PAULI_CHANNEL_1" else 2, args, [int(t) for t in parts[1:]]))
        elif op in ("HERALDED_ERASE", "HERALDED_PAULI_CHANNEL_1"):
            probs = [args[0] / 4.0] * 4 if op == "HERALDED_ERASE" else args
            for t in parts[1:]:
                q = int(t)
                site()
                out = []
                for w, psi, rec, nr in state["branches"]:
                    if 1.0 - sum(probs) > 0.0:
                        out.append((w * (1.0 - sum(probs)), psi, rec, nr + 1))
                    for pb, ax in zip(probs, [None, "X", "Y", "Z"]):
                        if pb > 0.0:
                            out.append((w * pb, _apply_ops(psi, n, [(q, ax)] if ax else []), rec | (1 << nr), nr + 1))
                state["branches"] = out
        elif op in ("M", "MR"):
            for t in parts[1:]:
                measure_z(int(t), op == "MR")
        elif op == "R":
            for t in parts[1:]:
                measure_z(int(t), True, record=False)
        elif op in _NOISE1 or op in ("DEPOLARIZE1", "DEPOLARIZE2"):
            noise(_pauli_terms(op, args[0], [int(t) for t in parts[1:]]))
        elif op in GATES1 or op in ("CX", "CNOT", "CZ"):
            state["branches"] = [(w, psi_apply(psi, n, op, [int(t) for t in parts[1:]]), rec, nr)
                                 for w, psi, rec, nr in state["branches"]]
        else:
            raise ValueError(f"unsupported instruction {op}")
    flush_chain()
    branches = state["branches"]
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
