"""Front end (parse -> lower -> optimize -> plan -> freeze) checks.

The planned program is pinned three ways: fixture shapes the reference
reports (max active width), invariants of noiseless stabilizer circuits
(detectors identically zero), and exact Born probabilities of small
circuits from a statevector (tests/statevector.py), sampled through the
oracle restatement of ticit's sampler.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from ticit_amd import TicitError, plan_text
from ticit_amd import workloads as w
from tests.statevector import expectation_values, final_distribution, record_values, tvd, with_probes

FIX = os.path.join(os.path.dirname(__file__), "golden", "circuits")


def read(name):
    with open(os.path.join(FIX, name)) as f:
        return f.read()


def test_blob_header_and_roundtrip_counts():
    blob = plan_text("H 0\nT 0\nH 0\nM 0\nOBSERVABLE_INCLUDE(0) rec[-1]\n")
    inf = O.info(blob)
    assert inf["n"] == 1 and inf["max_k"] == 1 and inf["nrecords"] == 1
    assert inf["nmeasure"] == 1 and inf["nobservables"] == 1


@pytest.mark.parametrize("name,max_k", [
    ("wide_compact.stim", 13),            # gpu/plan.rs:1595-1613
    ("deterministic_zero.stim", 0),
    ("dormant_random.stim", 0),
    ("active_random.stim", 1),
    ("two_active.stim", 2),
])
def test_fixture_active_widths(name, max_k):
    assert O.info(plan_text(read(name)))["max_k"] == max_k


def test_coherent_surface_width_matches_reference_note():
    # docs/impl-notes/perf/2026-08-10-gpu-correctness-and-comparison.md:140
    # reports coherent_d5_r5 at k = 22
    assert O.info(plan_text(read("coherent_surface_d5_r5_p1e-3_rz0p02.stim")))["max_k"] == 22


def test_deterministic_zero_fixture_rows():
    r = O.sample(plan_text(read("deterministic_zero.stim")), 64, 3, bit_packed=False)
    assert int(r["detectors"].sum()) == 0 and int(r["observable_ones"][0]) == 0


@pytest.mark.parametrize("d,rounds", [(3, 3), (5, 5)])
def test_noiseless_surface_code_detectors_are_zero(d, rounds):
    blob = plan_text(w.surface_code_t(d, rounds, 0.0, 0, 0))
    r = O.sample(blob, 3000, 17)
    assert int(np.unpackbits(r["detectors"], bitorder="little").sum()) == 0
    assert r["observable_ones"].tolist() == [0]


def test_noisy_surface_code_detector_rate_is_plausible():
    blob = plan_text(w.surface_code_t(5, 5, 1e-3, 0, 0))
    r = O.sample(blob, 20000, 1)
    rate = np.unpackbits(r["detectors"], axis=1, bitorder="little")[:, :120].mean()
    assert 0.004 < rate < 0.03


def small_circuits():
    out = []
    for i in range(6):
        rng = w.SplitMix64(w.block_seed(0xABCD, 7, i))
        n = 3 + i % 3
        out.append(w.random_clifford_t(n, 6 + i, 2 + i % 4, rng))
    out.append("H 0\nT 0\nH 0\nM 0\n")
    out.append("H 0 1\nT 0\nCX 0 1\nT_DAG 1\nH 0 1\nM 0 1\n")
    return out


@pytest.mark.parametrize("text", small_circuits())
def test_oracle_on_planned_program_matches_statevector(text):
    qubits, probs = final_distribution(text)
    blob = plan_text(text)
    shots = 40000
    r = O.sample(blob, shots, 123)
    vals = record_values(r["measurements"], len(qubits))
    counts = np.bincount(vals, minlength=probs.size).astype(float)
    # TVD of an exact sampler with 40000 shots over <= 32 outcomes stays well below 0.02
    assert tvd(counts, probs) < 0.02
    # outcomes of probability zero never appear
    assert counts[probs < 1e-12].sum() == 0


def test_config_programs_plan():
    for idx in (0, 1, 3):
        inf = O.info(plan_text(w.config_circuit(idx)))
        assert inf["nrecords"] > 0
    assert O.info(plan_text(w.config1_circuit()))["ndetectors"] == 120


def test_parse_errors_carry_kind():
    with pytest.raises(TicitError) as e:
        plan_text("H 0\nFOO 1\n")
    assert e.value.kind == "Unsupported"
    with pytest.raises(TicitError) as e:
        plan_text("M 0\nDETECTOR rec[-2]\n")
    assert e.value.kind == "Parse"
    with pytest.raises(TicitError) as e:
        plan_text("X_ERROR(1.5) 0\n")
    assert e.value.kind == "Parse"


def test_repeat_blocks_flatten():
    a = O.info(plan_text("REPEAT 3 {\nH 0\nM 0\n}\n"))
    assert a["nrecords"] == 3


def test_exp_val_allocates_one_slot_per_product():
    # src/circuit/parser.rs:1863-1872: slots 0..2 then 3; probes take no records
    inf = O.info(plan_text("EXP_VAL Z0*Z1 X0*X1 X2\nEXP_VAL Z2\n"))
    assert inf["nexpvals"] == 4
    assert inf["nrecords"] == 0
    with pytest.raises(TicitError) as e:
        plan_text("EXP_VAL\n")
    assert e.value.kind == "Parse" and "at least one Pauli product" in str(e.value)


def layered_probe_circuit(n: int, layers: int, rng: np.random.Generator) -> str:
    """H on every qubit, then layers of T/T_DAG, CX and H with X/Y/Z probes
    after each: most probes act on the active state with 0 < P < 1."""
    lines = ["H " + " ".join(map(str, range(n)))]
    for _ in range(layers):
        for q in range(n):
            r = rng.random()
            if r < 0.4:
                lines.append(("T " if rng.random() < 0.5 else "T_DAG ") + str(q))
            elif r < 0.55:
                lines.append("H " + str(q))
            elif r < 0.65:
                lines.append("S " + str(q))
        perm = rng.permutation(n)
        for a, b in zip(perm[0::2], perm[1::2]):
            if rng.random() < 0.5:
                lines.append(f"CX {int(a)} {int(b)}")
        prods = []
        for _ in range(int(rng.integers(2, 5))):
            size = 1 if rng.random() < 0.6 else min(n, 2)
            qs = rng.choice(n, size=size, replace=False)
            prods.append("*".join(("!" if rng.random() < 0.2 else "") + "XYXYZ"[int(rng.integers(5))] + str(int(q))
                                  for q in qs))
        lines.append("EXP_VAL " + " ".join(prods))
    lines.append("M " + " ".join(map(str, range(n))))
    return "\n".join(lines) + "\n"


def probe_circuits():
    out = []
    for i in range(6):
        rng = w.SplitMix64(w.block_seed(0xE7A1, 3, i))
        n = 2 + i % 4
        base = w.random_clifford_t(n, 5 + i, 1 + i % 4, rng)
        out.append(with_probes(base, np.random.default_rng(100 + i), 3 + i % 3))
    for i in range(6):
        out.append(layered_probe_circuit(2 + i % 4, 4 + i, np.random.default_rng(200 + i)))
    return out


@pytest.mark.parametrize("text", probe_circuits())
def test_expectation_probes_match_statevector(text):
    """Probes before any measurement read <psi|P|psi> of the deterministic
    state, on every shot; the planner keeps them where they were written
    (src/pending_optimizer.rs:230-242)."""
    want = expectation_values(text)
    blob = plan_text(text)
    assert O.info(blob)["nexpvals"] == want.size
    r = O.sample(blob, 64, 5, bit_packed=True)
    assert r["exp_vals"].shape == (64, want.size)
    if text.startswith("H 0 1"):  # the layered family: fractional probes exercise 0 < P < 1
        assert np.sum(np.abs(np.abs(want) - np.round(np.abs(want))) > 1e-3) >= 2
    assert np.abs(r["exp_vals"] - want[None, :]).max() <= 1e-12


def test_probes_do_not_change_the_sampled_distribution():
    """A probe is non-destructive: the final record distribution is the
    probe-free circuit's exact one."""
    text = small_circuits()[3]
    qubits, probs = final_distribution(text)
    probed = with_probes(text, np.random.default_rng(9), 4)
    r = O.sample(plan_text(probed), 40000, 77)
    vals = record_values(r["measurements"], len(qubits))
    counts = np.bincount(vals, minlength=probs.size).astype(float)
    assert tvd(counts, probs) < 0.02
