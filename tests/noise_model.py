"""Independent restatement of ticit's exogenous noise draws (test-only).

Written from the reference's own text, not from the oracle, so that the
oracle's C++ walk (and through it the device walk) is checked against a
second reading of the published algorithm:

  * SplitMix64, rand_float, geometric gaps, categorical rows, block_seed
    (src/random.rs:13-24, 26-50, 55-80);
  * the low-probability skip walk (src/sampler/exogenous.rs:150-175), the
    8-coin biased-bit generator (exogenous.rs:185-266), the rare categorical
    walk (exogenous.rs:304-354) and the dense low-probability group walk
    (exogenous.rs:410-440), in the family order of
    resample_prepared_exogenous_packed_in_place (exogenous.rs:536-625), each
    run with shots = 1 because the hip backend gives every shot its own noise
    stream block_seed(0x7eed0000, seed, shot) (src/sampler/prepared.rs:753-778
    at sample_chunk_shots = 1);
  * the lowering tables: X/Y/Z_ERROR mint one Bernoulli condition per target
    (src/circuit/lowering.rs:690-700); DEPOLARIZE1 is a 4-row categorical
    over (x, z) with rows I, Z, X, Y and weights 1-p, p/3, p/3, p/3
    (lowering.rs:143-159); DEPOLARIZE2 a 16-row categorical over
    (x1, z1, x2, z2), row r>0 holding the bits of r with x1 most
    significant, weights 1-p then p/15 (lowering.rs:161-186);
  * the planner's split into families (src/planner.rs:1301-1338,1344-1383,
    1455-1479): a categorical whose event probability 1 - P(row 0) is below
    0.02 joins a rare group (event rows renormalized by the event
    probability), Bernoulli conditions below 0.02 join a low-probability
    group per exact probability, and every family keeps creation order
    (Bernoulli conditions in ascending condition order).

A channel is described as a list of "sites", one per noise call in program
order; `sample_shot` returns each site's drawn Pauli bits for one shot.
"""
from __future__ import annotations

import math

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
EXO_BASE = 0x7EED0000
LOW_P = 0.02
DBL_MIN = 2.2250738585072014e-308
I32_MAX = 2147483647.0


def next_u64(state: list) -> int:
    state[0] = (state[0] + GOLDEN) & MASK64
    z = state[0]
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def rand_float(state: list) -> float:
    return (next_u64(state) >> 11) * 2.0 ** -53


def block_seed(base: int, seed: int, block: int) -> int:
    return (base ^ ((GOLDEN * (block + 1)) & MASK64) ^ ((0xBF58476D1CE4E5B9 * (seed + 1)) & MASK64)) & MASK64


def gap(state: list, denom: float) -> float:
    u = max(rand_float(state), DBL_MIN)
    g = math.floor(math.log(u) / denom)
    if not math.isfinite(g) or g >= I32_MAX:
        return I32_MAX
    return float(g)


def categorical_row(state: list, probs: list) -> int:
    r = rand_float(state)
    cum = 0.0
    for i, p in enumerate(probs):
        cum += p
        if r <= cum:
            return i
    return len(probs) - 1


def low_probability_bit(state: list, p: float) -> bool:
    """or_low_probability_bits_packed over one shot: one live draw."""
    if p <= 0.0:
        return False
    denom = math.log1p(-p)
    draw = 0
    hit = False
    while True:
        g = int(gap(state, denom))
        if g >= 1 - draw:
            return hit
        draw += g
        hit = True
        draw += 1


def biased_bit(state: list, probability: float) -> bool:
    """generate_packed_biased_bits for one shot (bit 0 of the row)."""
    if probability <= 0.0:
        return False
    if probability >= 1.0:
        return True
    invert = probability > 0.5
    p = 1.0 - probability if invert else probability
    if p <= 0.0:
        return invert
    if p == 0.5:
        bit = bool(next_u64(state) & 1)
    elif p < LOW_P:
        bit = low_probability_bit(state, p)
    else:
        top = int(p * 256.0)
        top = top if top < 128 else 127
        p_trunc = top / 256.0
        alive = next_u64(state)
        result = 0
        for b in range(6, -1, -1):
            shoot = next_u64(state)
            if (top >> b) & 1:
                result |= shoot & alive
            alive &= ~shoot & MASK64
        bit = bool(result & 1)
        left = p - p_trunc
        if left > 0.0:
            bit = low_probability_bit(state, left / (1.0 - p_trunc)) or bit
    return (not bit) if invert else bit


def depolarize1_table(p: float):
    """(assignments as bit lists [x, z], probabilities) in row order."""
    third = p / 3.0
    return [[0, 0], [0, 1], [1, 0], [1, 1]], [1.0 - p, third, third, third]


def depolarize2_table(p: float):
    rows = [[0, 0, 0, 0]]
    for r in range(1, 16):
        rows.append([(r >> 3) & 1, (r >> 2) & 1, (r >> 1) & 1, r & 1])  # x1, z1, x2, z2
    probs = [p / 15.0] * 16
    probs[0] = 1.0 - p
    return rows, probs


class Plan:
    """The exogenous families of a program whose noise calls are `sites`.

    Each site is ("X"|"Y"|"Z", p) for a one-qubit Pauli error (one Bernoulli
    condition; Y sets both bits), or ("DEP1", p) / ("DEP2", p). Sites are in
    program order, which is condition order.
    """

    def __init__(self, sites, dep2_rows=None, swap_set_and_shot=False, stay_on_hit=False):
        self.sites = list(sites)
        # planted mistakes for sensitivity checks: the set of a rare hit
        # taken from draw / nsets (the shot index) instead of draw % nsets;
        # a skip walk that does not step past a hit (no draw += 1)
        self.swap_set_and_shot = swap_set_and_shot
        self.advance = 0 if stay_on_hit else 1
        self.cats = []      # (site, rows, probs)
        self.rares = []     # dict(event_p, rows, probs, event_rows, event_probs, sets=[site...])
        self.berns = []     # (site, p)
        self.lowps = []     # dict(p, sites=[...])
        for i, (kind, p) in enumerate(self.sites):
            if kind in ("X", "Y", "Z"):
                if p < LOW_P:
                    for g in self.lowps:
                        if g["p"] == p:
                            g["sites"].append(i)
                            break
                    else:
                        self.lowps.append({"p": p, "sites": [i]})
                else:
                    self.berns.append((i, p))
                continue
            rows, probs = depolarize1_table(p) if kind == "DEP1" else (dep2_rows or depolarize2_table)(p)
            event_p = 1.0 - probs[0]
            if event_p < LOW_P:
                inv = 1.0 / event_p if event_p > 0.0 else 0.0
                ev_rows = [r for r in range(1, len(rows)) if probs[r] > 0.0] if event_p > 0.0 else []
                ev_probs = [probs[r] * inv for r in ev_rows]
                for g in self.rares:
                    if g["event_p"] == event_p and g["rows"] == rows and g["probs"] == probs:
                        g["sets"].append(i)
                        break
                else:
                    self.rares.append({"event_p": event_p, "rows": rows, "probs": probs,
                                       "event_rows": ev_rows, "event_probs": ev_probs, "sets": [i]})
            else:
                self.cats.append((i, rows, probs))

    def width(self, i: int) -> int:
        kind = self.sites[i][0]
        return 4 if kind == "DEP2" else 2

    def sample_shot(self, seed: int, shot: int):
        """Per site, its Pauli bits for this shot: [x, z] (one-qubit sites)
        or [x1, z1, x2, z2] (DEPOLARIZE2)."""
        out = [[0] * self.width(i) for i in range(len(self.sites))]
        st = [block_seed(EXO_BASE, seed, shot)]
        for i, rows, probs in self.cats:
            out[i] = list(rows[categorical_row(st, probs)])
        for g in self.rares:
            if g["event_p"] <= 0.0:
                continue
            nsets = len(g["sets"])
            denom = math.log1p(-g["event_p"])
            draw = 0
            while True:
                d = int(gap(st, denom))
                if d >= nsets - draw:
                    break
                draw += d
                set_idx = draw // nsets if self.swap_set_and_shot else draw % nsets  # shot = draw // nsets = 0
                row = g["event_rows"][categorical_row(st, g["event_probs"])]
                site = g["sets"][set_idx]
                out[site] = [a ^ b for a, b in zip(out[site], g["rows"][row])]
                draw += self.advance
        for i, p in self.berns:
            if biased_bit(st, p):
                out[i] = pauli_bits(self.sites[i][0])
        for g in self.lowps:
            if g["p"] <= 0.0:
                continue
            n = len(g["sites"])
            denom = math.log1p(-g["p"])
            draw = 0
            while True:
                d = int(gap(st, denom))
                if d >= n - draw:
                    break
                draw += d
                site = g["sites"][draw % n]
                out[site] = [a ^ b for a, b in zip(out[site], pauli_bits(self.sites[site][0]))]
                draw += self.advance
        return out


def pauli_bits(axis: str) -> list:
    return {"X": [1, 0], "Y": [1, 1], "Z": [0, 1]}[axis]


def bell_decode_circuit(sites) -> tuple[str, list]:
    """A noiseless-zero circuit whose records expose every site's bits.

    Every one-qubit site acts on qubit a of a Bell pair (a, a'); a
    DEPOLARIZE2 site on the first qubits of two pairs. After the pairs are
    unprepared (CX a a', H a) an X on a flips M a', a Z flips M a, so the
    record of a holds z and that of a' holds x. Returns the text and, per
    site, the record indices of its bits in the order sample_shot gives
    them.
    """
    pairs = []
    noise = []
    for kind, p in sites:
        if kind == "DEP2":
            a, b = len(pairs), len(pairs) + 1
            pairs += [a, b]
            noise.append((kind, p, [a, b]))
        else:
            a = len(pairs)
            pairs.append(a)
            noise.append((kind, p, [a]))
    n = 2 * len(pairs)
    first = " ".join(str(2 * k) for k in range(len(pairs)))
    cx = " ".join(f"{2 * k} {2 * k + 1}" for k in range(len(pairs)))
    lines = [f"H {first}", f"CX {cx}"]
    # consecutive sites of one channel share a line: its targets are lowered
    # one call per target (pair), in target order
    j = 0
    while j < len(noise):
        kind, p, _ = noise[j]
        name = {"X": "X_ERROR", "Y": "Y_ERROR", "Z": "Z_ERROR", "DEP1": "DEPOLARIZE1", "DEP2": "DEPOLARIZE2"}[kind]
        targets = []
        while j < len(noise) and noise[j][:2] == (kind, p):
            targets += [str(2 * k) for k in noise[j][2]]
            j += 1
        lines.append(f"{name}({p!r}) " + " ".join(targets))
    lines += [f"CX {cx}", f"H {first}", "M " + " ".join(str(q) for q in range(n))]
    recs = []
    for kind, p, prs in noise:
        idx = []
        for k in prs:
            idx += [2 * k + 1, 2 * k]  # x from M a', z from M a
        recs.append(idx)
    return "\n".join(lines) + "\n", recs


def predicted_records(sites, seed: int, shots: int, dep2_rows=None, swap_set_and_shot=False, stay_on_hit=False):
    """[shots][nrec] uint8 records the bell_decode_circuit of `sites` must
    produce at `seed`, from the restated walks alone."""
    import numpy as np
    text, recs = bell_decode_circuit(sites)
    nrec = 2 * sum(2 if k == "DEP2" else 1 for k, _ in sites)  # two records per Bell pair
    plan = Plan(sites, dep2_rows, swap_set_and_shot, stay_on_hit)
    out = np.zeros((shots, nrec), dtype=np.uint8)
    for s in range(shots):
        bits = plan.sample_shot(seed, s)
        for i, idx in enumerate(recs):
            for b, r in zip(bits[i], idx):
                out[s, r] = b
    return out
