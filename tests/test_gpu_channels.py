"""The front end's newer instructions on the device (gpu-marked).

The cases of tests/test_frontend_channels.py on every tier: seeded rows
bit-exact with the oracle at the per-shot stream contract, and record
distributions against the exact statevector mixture by chi-square. Then the
reference's own CCZ fixture (testdata/circuits/ccz/d05_p1e-3.clifft: 834
qubits, MPP, CY, MY with readout noise, E(0.125) markers, 6992 detectors and
28 EXP_VAL probes), normalized to its reference sample as the reference's
benchmark runs it: rows, counts and expectation values against the oracle.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.statevector import chi2_pvalue, record_distribution, record_values
from tests.test_frontend_channels import CHANNEL_CASES

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "circuits")


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
@pytest.mark.parametrize("name", sorted(CHANNEL_CASES))
def test_new_instructions_on_every_tier(hip, name, tier):
    text = CHANNEL_CASES[name]
    c = hip.Circuit(text)
    s = c.compile(tier=tier)
    shots = 2000
    r = s.sample(shots, seed=13, bit_packed=True)
    o = O.sample_per_shot(c.blob, shots, 13, bit_packed=True)
    assert np.array_equal(r.measurements, o["measurements"])
    dist = record_distribution(text)
    nrec = int(np.log2(dist.size))
    big = s.sample(200000, seed=5, bit_packed=True)
    counts = np.bincount(record_values(big.measurements, nrec), minlength=dist.size)
    pv = chi2_pvalue(counts, dist)
    print(name, tier, "max_k", s.max_active_qubits, "chi2 p-value", pv, flush=True)
    assert counts[dist < 1e-12].sum() == 0
    assert pv > 1e-4


def test_ccz_d05_fixture_matches_the_oracle(hip):
    with open(os.path.join(GOLDEN, "ccz_d05_p1e-3.clifft")) as f:
        c = hip.Circuit(f.read())
    s = c.compile(normalize_syndromes=True)
    ref_d, ref_o = O.reference_sample(c.blob)
    assert s.reference_sample().detectors == [bool(v) for v in ref_d]
    shots = 400
    r = s.sample(shots, seed=17, bit_packed=True)
    O.set_expected(ref_d, ref_o)
    try:
        o = O.sample_per_shot(c.blob, shots, 17, bit_packed=True)
    finally:
        O.set_expected(None, None)
    print("ccz d05 tier", s.tier, "max_k", s.max_active_qubits, "detector ones", int(np.unpackbits(r.detectors).sum()),
          flush=True)
    assert (r.shots, r.discarded, r.accepted) == tuple(int(v) for v in o["counts"][:3])
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(r, key), o[key]), key
    assert r.exp_vals.shape == (shots, 28)
    assert np.max(np.abs(r.exp_vals - o["exp_vals"])) <= 1e-12
