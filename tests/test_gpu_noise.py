"""The device's exogenous noise draws, pinned without the oracle (gpu-marked).

The same three checks as tests/test_noise_pinning.py, on device rows:
bit-exact records against the independent restatement of the reference's
walks and lowering tables (tests/noise_model.py) for every channel at
p = 1e-3, 0.015, 0.05, 0.3 and for a program mixing every family on each
tier; per-Pauli flip rates against p, p/3 and p/15 by chi-square at 4e5
shots; and noisy circuits with mid-circuit measurements against the exact
mixture (tests/statevector.py:record_distribution) on each tier.
"""
import numpy as np
import pytest

from tests import noise_model as N
from tests.statevector import chi2_pvalue, noisy_midcircuit_circuits, record_distribution, record_values
from tests.test_noise_pinning import CHANNELS, MIXED, PROBS, closed_form, pauli_counts, sites_of

pytestmark = pytest.mark.gpu


def device_records(hip, sites, shots, seed, tier="auto"):
    text, _ = N.bell_decode_circuit(sites)
    s = hip.Circuit(text).compile(**({} if tier == "auto" else {"tier": tier}))
    r = s.sample(shots, seed=seed, bit_packed=False)
    return r.measurements


@pytest.mark.parametrize("p", PROBS)
@pytest.mark.parametrize("kind", sorted(CHANNELS))
def test_device_rows_equal_the_restated_walks(hip, kind, p):
    sites = sites_of(kind, p)
    shots = 3000
    got = device_records(hip, sites, shots, 5)
    print(kind, p, "ones", int(got.sum()), flush=True)
    assert np.array_equal(got, N.predicted_records(sites, 5, shots))


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
def test_every_family_on_every_tier(hip, tier):
    shots = 2000
    got = device_records(hip, MIXED, shots, 9, tier)
    assert np.array_equal(got, N.predicted_records(MIXED, 9, shots))


@pytest.mark.parametrize("p", PROBS)
@pytest.mark.parametrize("kind", sorted(CHANNELS))
def test_device_flip_rates_follow_the_closed_forms(hip, kind, p):
    sites = sites_of(kind, p)
    rows = device_records(hip, sites, 400000, 77)
    counts = pauli_counts(rows, sites)
    probs = closed_form(kind, p)
    assert counts[probs == 0].sum() == 0
    pv = chi2_pvalue(counts, probs)
    print(kind, p, "chi2 p-value", pv, flush=True)
    assert pv > 1e-4, (counts, probs * counts.sum())


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
@pytest.mark.parametrize("index", range(6))
def test_noisy_midcircuit_records_follow_the_exact_mixture(hip, tier, index):
    text = noisy_midcircuit_circuits()[index]
    dist = record_distribution(text)
    nrec = int(np.log2(dist.size))
    shots = 200000
    s = hip.Circuit(text).compile(tier=tier)
    r = s.sample(shots, seed=3, bit_packed=True)
    counts = np.bincount(record_values(r.measurements, nrec), minlength=dist.size)
    assert counts[dist < 1e-12].sum() == 0
    pv = chi2_pvalue(counts, dist)
    print(tier, index, "max_k", s.max_active_qubits, "chi2 p-value", pv, flush=True)
    assert pv > 1e-4
