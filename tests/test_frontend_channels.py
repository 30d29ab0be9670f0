"""The front end's newer instructions against exact mixtures (CPU).

E / CORRELATED_ERROR / ELSE_CORRELATED_ERROR chains, MPP / MXX / MYY /
MZZ, SPP / SPP_DAG / R_PAULI, PAULI_CHANNEL_1 / 2, HERALDED_ERASE /
HERALDED_PAULI_CHANNEL_1 and MPAD (src/circuit/parser.rs:987-1021,
1440-1590; src/circuit/lowering.rs:195-366, 585-594), planned by the C++
front end and sampled by the oracle, against the record distribution of a
direct statevector mixture (tests/statevector.py:record_distribution), by
chi-square; and the parse errors the reference raises for them.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.statevector import chi2_pvalue, record_distribution, record_values
from ticit_amd import TicitError, plan_text

# PAULI_CHANNEL_2: fifteen distinct probabilities, 0.1 in total
P2 = ",".join(f"{0.1 * (i + 1) / 120:.6g}" for i in range(15))

CHANNEL_CASES = {
    "correlated_chain": "H 0 1\nT 0\nCX 0 2\nE(0.2) X0 Z1\nELSE_CORRELATED_ERROR(0.3) Y1 X2\n"
                        "ELSE_CORRELATED_ERROR(0.5) Z0\nH 0 1 2\nM 0 1 2\n",
    "two_chains": "H 0\nT 0\nCORRELATED_ERROR(0.25) X1\nE(0.4) X0 X1\nH 0\nM 0 1\n",
    "pair_measurements": "H 0 1 2\nT 0\nCX 0 1\nT 1\nMXX 0 1\nMZZ !1 2\nMYY 0 2\nM 0 1 2\n",
    "mpp_products": "H 0 1\nT 0\nCX 0 2\nMPP X0*Z1 !Y1*X2 Z0*Z1*Z2\nM 0 1 2\n",
    "pauli_rotations": "H 0 1\nT 0\nSPP X0*Z1\nSPP_DAG Y1\nR_PAULI(0.25) X0*X1\nH 0\nM 0 1\n",
    "pauli_channels": f"H 0\nT 0\nPAULI_CHANNEL_1(0.1,0.2,0.05) 0\nCX 0 1\nPAULI_CHANNEL_2({P2}) 0 1\nH 1\nM 0 1\n",
    "heralds": "H 0\nT 0\nCX 0 1\nHERALDED_ERASE(0.3) 1\nHERALDED_PAULI_CHANNEL_1(0.1,0.2,0.15,0.05) 0\nH 0\nM 0 1\n",
    "mpad": "MPAD 1 0 !1\nH 0\nT 0\nH 0\nM 0\nMPAD(0.2) 0\nX_ERROR(0.1) 0\nM 0\n",
}


@pytest.mark.parametrize("name", sorted(CHANNEL_CASES))
def test_oracle_on_new_instructions_follows_the_exact_mixture(name):
    text = CHANNEL_CASES[name]
    dist = record_distribution(text)
    nrec = int(np.log2(dist.size))
    shots = 60000
    r = O.sample_per_shot(plan_text(text), shots, 11, bit_packed=True)
    counts = np.bincount(record_values(r["measurements"], nrec), minlength=dist.size)
    assert counts[dist < 1e-12].sum() == 0
    assert chi2_pvalue(counts, dist) > 1e-4


def test_zero_probability_chain_changes_nothing():
    """E(0) (the reference's benchmark rewrite of its CCZ fixtures) keeps
    the noiseless distribution."""
    base = "H 0\nT 0\nH 0\nM 0\n"
    with_chain = "H 0\nT 0\nE(0) Z0\nELSE_CORRELATED_ERROR(0) X0\nH 0\nM 0\n"
    assert np.allclose(record_distribution(base), record_distribution(with_chain))
    a = O.sample_per_shot(plan_text(base), 5000, 3)["measurements"]
    b = O.sample_per_shot(plan_text(with_chain), 5000, 3)["measurements"]
    assert abs(float(a.mean()) - float(b.mean())) < 0.03


@pytest.mark.parametrize("text,needle", [
    ("ELSE_CORRELATED_ERROR(0.1) X0\nM 0\n", "must follow CORRELATED_ERROR"),
    ("MXX 0 1 2\n", "paired targets"),
    ("MZZ 1 1\n", "distinct qubits"),
    ("PAULI_CHANNEL_1(0.1,0.2) 0\n", "does not match gate arity"),
    ("MPAD 2\n", "must be 0 or 1"),
    ("E(0.1) X0*Z1\n", "combiner targets"),
    ("HERALDED_PAULI_CHANNEL_1(0.1,0.2) 0\n", "four probabilities"),
    ("I_ERROR(0.7,0.6) 0\n", "sum to at most 1"),
])
def test_new_instruction_parse_errors(text, needle):
    with pytest.raises(TicitError) as e:
        plan_text(text)
    assert needle in str(e.value)
