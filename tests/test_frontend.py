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
from tests.statevector import final_distribution, record_values, tvd

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
