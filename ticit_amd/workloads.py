"""Seeded synthetic circuits for the five benchmark configurations.

Every generator draws from SplitMix64 (the generator of ticit's
src/random.rs:13-24) seeded with ``block_seed(0xC0F1_0000, config, index)``
(src/random.rs:77-80) and emits ``.ticit`` text, so the CPU oracle and the
HIP backend always see the same circuit.

Configurations:
  * C0 ``random_clifford_t`` — 16 qubits, 40 layers, 12 T/T_DAG, final M;
  * C1 ``surface_code_t`` — rotated d=5 Z-memory, 5 rounds, DEPOLARIZE1/2
    p=1e-3, X_ERROR p=1e-3 before measurements, two T per round;
  * C2 ``batch_circuits`` — 256 random circuits (8-40 qubits, 4-20 T,
    DEPOLARIZE1 5e-4 after every one-qubit gate), log-uniform shot counts;
  * C3 ``dense_stress`` — 30 qubits, 60 layers, 24 T, mid-circuit M on 8
    qubits after layers 15/30/45, final M;
  * C4 ``planner_corpus`` — varied circuits from the C0/C3 generator.
"""
from __future__ import annotations

import math

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
CONFIG_SEED_BASE = 0xC0F1_0000


def block_seed(base: int, seed: int, block_index: int) -> int:
    """ticit's seed derivation for one deterministic work unit."""
    return (base ^ ((GOLDEN * (block_index + 1)) & MASK64)
            ^ ((0xBF58476D1CE4E5B9 * (seed + 1)) & MASK64)) & MASK64


class SplitMix64:
    def __init__(self, state: int):
        self.state = state & MASK64

    def next(self) -> int:
        self.state = (self.state + GOLDEN) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (2.0 ** -53)

    def below(self, n: int) -> int:
        return self.next() % n


ONE_QUBIT_GATES = ("H", "S", "S_DAG", "X", "Z")


def random_clifford_t(n: int, layers: int, t_count: int, rng: SplitMix64,
                      depolarize1: float = 0.0,
                      mid_measure: dict | None = None) -> str:
    """The C0 layer generator.

    Per layer: every qubit takes a gate from {H,S,S_DAG,X,Z} with probability
    1/2; then T/T_DAG at the seeded (layer, qubit) slots; then n/4 disjoint
    CX pairs of random orientation on n/2 random qubits. ``depolarize1``
    appends DEPOLARIZE1 after every one-qubit gate (C2). ``mid_measure`` maps
    a layer index to the qubits measured after that layer (C3).
    """
    slots = n * layers
    if t_count > slots:
        raise ValueError("more T gates than (layer, qubit) slots")
    t_at: dict[int, list[tuple[int, str]]] = {}
    used: set[int] = set()
    while len(used) < t_count:
        slot = rng.below(slots)
        if slot in used:
            continue
        used.add(slot)
        gate = "T" if rng.next() & 1 else "T_DAG"
        t_at.setdefault(slot // n, []).append((slot % n, gate))
    lines: list[str] = []
    for layer in range(layers):
        one_q: list[int] = []
        for q in range(n):
            if rng.uniform() < 0.5:
                gate = ONE_QUBIT_GATES[rng.below(len(ONE_QUBIT_GATES))]
                lines.append(f"{gate} {q}")
                one_q.append(q)
        if depolarize1 > 0.0 and one_q:
            lines.append(f"DEPOLARIZE1({depolarize1}) " + " ".join(map(str, one_q)))
        for q, gate in sorted(t_at.get(layer, [])):
            lines.append(f"{gate} {q}")
            if depolarize1 > 0.0:
                lines.append(f"DEPOLARIZE1({depolarize1}) {q}")
        order = list(range(n))
        picks = 2 * (n // 4)
        for i in range(picks):
            j = i + rng.below(n - i)
            order[i], order[j] = order[j], order[i]
        pairs = []
        for i in range(0, picks, 2):
            a, b = order[i], order[i + 1]
            if rng.next() & 1:
                a, b = b, a
            pairs.append(f"{a} {b}")
        if pairs:
            lines.append("CX " + " ".join(pairs))
        if mid_measure and layer in mid_measure:
            lines.append("M " + " ".join(map(str, mid_measure[layer])))
        lines.append("TICK")
    lines.append("M " + " ".join(map(str, range(n))))
    return "\n".join(lines) + "\n"


def config0_circuit() -> str:
    rng = SplitMix64(block_seed(CONFIG_SEED_BASE, 0, 0))
    return random_clifford_t(16, 40, 12, rng)


def surface_code_t(distance: int = 5, rounds: int = 5, p: float = 1e-3,
                   t_per_round: int = 2, index: int = 0) -> str:
    """Rotated surface-code Z-memory with T gates injected on data qubits.

    Data qubits sit at odd (x, y) in [1, 2d-1]; measure qubits at even
    (x, y) in [0, 2d] with a checkerboard of X/Z type (interior plaquettes
    plus weight-2 X checks on the top/bottom and Z checks on the left/right
    boundaries). CX order per check follows the standard hook-safe schedule.
    Each round: T on ``t_per_round`` seeded data qubits (+DEPOLARIZE1), H on X
    checks (+DEPOLARIZE1), four CX layers (+DEPOLARIZE2), H (+DEPOLARIZE1),
    X_ERROR on the measure qubits, MR. Detectors compare each check with its
    previous round (Z checks also against the initial |0> data in round 1);
    the final data readout closes the Z checks and includes logical Z.
    """
    d = distance
    rng = SplitMix64(block_seed(CONFIG_SEED_BASE, 1, index))
    data = [(x, y) for y in range(1, 2 * d, 2) for x in range(1, 2 * d, 2)]
    checks = []
    for y in range(0, 2 * d + 1, 2):
        for x in range(0, 2 * d + 1, 2):
            kind = "X" if ((x // 2) + (y // 2)) % 2 == 0 else "Z"
            nb = [(x + dx, y + dy) for dx in (-1, 1) for dy in (-1, 1)
                  if 0 < x + dx < 2 * d and 0 < y + dy < 2 * d]
            if len(nb) == 4:
                checks.append((x, y, kind))
            elif len(nb) == 2:
                on_tb = y in (0, 2 * d)
                on_lr = x in (0, 2 * d)
                if on_tb and not on_lr and kind == "X":
                    checks.append((x, y, kind))
                elif on_lr and not on_tb and kind == "Z":
                    checks.append((x, y, kind))
    index_of = {}
    for i, (x, y) in enumerate(data):
        index_of[(x, y)] = i
    for j, (x, y, _) in enumerate(checks):
        index_of[(x, y)] = len(data) + j
    xs = [index_of[(x, y)] for (x, y, k) in checks if k == "X"]
    zs = [index_of[(x, y)] for (x, y, k) in checks if k == "Z"]
    anc = [index_of[(x, y)] for (x, y, _) in checks]
    data_idx = [index_of[q] for q in data]
    x_order = [(1, 1), (1, -1), (-1, 1), (-1, -1)]
    z_order = [(1, 1), (-1, 1), (1, -1), (-1, -1)]
    layers = []
    for step in range(4):
        pairs = []
        for (x, y, k) in checks:
            dx, dy = (x_order if k == "X" else z_order)[step]
            nb = (x + dx, y + dy)
            if nb not in index_of or nb not in set(data):
                continue
            a, q = index_of[(x, y)], index_of[nb]
            pairs.append((a, q) if k == "X" else (q, a))
        layers.append(pairs)
    fmt = lambda qs: " ".join(map(str, qs))
    lines = []
    for (x, y) in data:
        lines.append(f"QUBIT_COORDS({x}, {y}) {index_of[(x, y)]}")
    for (x, y, _) in checks:
        lines.append(f"QUBIT_COORDS({x}, {y}) {index_of[(x, y)]}")
    lines.append("R " + fmt(data_idx + anc))
    lines.append("TICK")
    nanc = len(anc)
    for r in range(rounds):
        picks = []
        while len(picks) < t_per_round:
            q = data_idx[rng.below(len(data_idx))]
            if q not in picks:
                picks.append(q)
        lines.append("T " + fmt(picks))
        lines.append(f"DEPOLARIZE1({p}) " + fmt(picks))
        lines.append("H " + fmt(xs))
        lines.append(f"DEPOLARIZE1({p}) " + fmt(xs))
        lines.append("TICK")
        for pairs in layers:
            flat = [q for pr in pairs for q in pr]
            lines.append("CX " + fmt(flat))
            lines.append(f"DEPOLARIZE2({p}) " + fmt(flat))
            lines.append("TICK")
        lines.append("H " + fmt(xs))
        lines.append(f"DEPOLARIZE1({p}) " + fmt(xs))
        lines.append(f"X_ERROR({p}) " + fmt(anc))
        lines.append("MR " + fmt(anc))
        for j, a in enumerate(anc):
            kind = checks[j][2]
            cur = f"rec[{-nanc + j}]"
            if r == 0:
                if kind == "Z":
                    lines.append(f"DETECTOR {cur}")
            else:
                lines.append(f"DETECTOR {cur} rec[{-2 * nanc + j}]")
        lines.append("TICK")
    lines.append(f"X_ERROR({p}) " + fmt(data_idx))
    lines.append("M " + fmt(data_idx))
    ndata = len(data_idx)
    pos = {q: i for i, q in enumerate(data_idx)}
    for j, (x, y, kind) in enumerate(checks):
        if kind != "Z":
            continue
        recs = [f"rec[{-ndata + pos[index_of[(x + dx, y + dy)]]}]"
                for dx in (-1, 1) for dy in (-1, 1)
                if (x + dx, y + dy) in index_of and index_of[(x + dx, y + dy)] in pos]
        recs.append(f"rec[{-ndata - nanc + j}]")
        lines.append("DETECTOR " + " ".join(recs))
    # logical Z: the data row y = 1, which runs between the two Z-type
    # (left/right) boundaries and so commutes with every X check
    col = [pos[index_of[(x, 1)]] for x in range(1, 2 * d, 2)]
    lines.append("OBSERVABLE_INCLUDE(0) " + " ".join(f"rec[{-ndata + i}]" for i in col))
    return "\n".join(lines) + "\n"


def config1_circuit() -> str:
    return surface_code_t(5, 5, 1e-3, 2, 0)


def config2_batch(count: int = 256):
    """(circuit text, shots) pairs of the batching configuration."""
    out = []
    for i in range(count):
        rng = SplitMix64(block_seed(CONFIG_SEED_BASE, 2, i))
        n = 8 + rng.below(33)
        t = 4 + rng.below(17)
        u = rng.uniform()
        shots = int(round(10 ** (4 + 2 * u)))
        out.append((random_clifford_t(n, 40, t, rng, depolarize1=5e-4), shots))
    return out


def config3_circuit() -> str:
    rng = SplitMix64(block_seed(CONFIG_SEED_BASE, 3, 0))
    mids = {}
    for layer in (14, 29, 44):  # after layers 15, 30 and 45 (1-based)
        qs = list(range(30))
        for i in range(8):
            j = i + rng.below(30 - i)
            qs[i], qs[j] = qs[j], qs[i]
        mids[layer] = sorted(qs[:8])
    return random_clifford_t(30, 60, 24, rng, mid_measure=mids)


def planner_corpus_circuit(index: int) -> str:
    """C4 corpus: circuits of the C0 layer generator with varied width, depth
    and T count, half of them with C3-style mid-circuit measurement layers.
    Sized so that most programs reach the active widths (max_k >= 8) where
    dense and tensor-component execution are both available."""
    rng = SplitMix64(block_seed(CONFIG_SEED_BASE, 4, index))
    n = 10 + rng.below(19)          # 10..28 qubits
    layers = 10 + rng.below(51)     # 10..60 layers
    t = 8 + rng.below(17)           # 8..24 T gates
    mids = {}
    if rng.next() & 1:
        every = 10 + rng.below(11)
        for layer in range(every - 1, layers - 1, every):
            m = 1 + rng.below(max(1, n // 4))
            qs = list(range(n))
            for i in range(m):
                j = i + rng.below(n - i)
                qs[i], qs[j] = qs[j], qs[i]
            mids[layer] = sorted(qs[:m])
    return random_clifford_t(n, layers, t, rng, mid_measure=mids)


CONFIGS = {
    0: dict(name="random_clifford_t_n16_l40_t12", shots=100_000),
    1: dict(name="surface_d5_r5_p1e-3_t10", shots=1_000_000),
    3: dict(name="dense_n30_l60_t24", shots=100_000),
}


def config_circuit(index: int) -> str:
    if index == 0:
        return config0_circuit()
    if index == 1:
        return config1_circuit()
    if index == 3:
        return config3_circuit()
    raise ValueError(f"config {index} is not a single circuit")


def log_uniform_shots(u: float) -> int:
    return int(round(10 ** (4 + 2 * u)))


__all__ = [
    "SplitMix64", "block_seed", "random_clifford_t", "surface_code_t",
    "config0_circuit", "config1_circuit", "config2_batch", "config3_circuit",
    "planner_corpus_circuit", "config_circuit", "CONFIGS", "math",
]
