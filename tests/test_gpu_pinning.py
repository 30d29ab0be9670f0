"""Pinned measurement parities on the GPU (gpu-marked), against the CPU
oracle. A pinned branch takes the XOR of earlier branches instead of a draw
(src/sampler/pinning.rs, src/sampler/batch/active.rs:535-554); rows are
bit-exact with the oracle at the per-shot stream contract, a biased pinned
branch fails sampling ("not one half") and a decided constraint fails
creation ("deterministic")."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle.pinning import decode_instructions, plan_pinned, PinError
from ticit_amd import workloads as w

pytestmark = pytest.mark.gpu


def surface_pins(blob, want=4):
    """Constraints on single surface-code records whose pinned branch is a
    dormant coin (always fair)."""
    ins = decode_instructions(blob)
    pins = []
    for r in range(24, 120):
        try:
            f = plan_pinned(blob, pins + [([r], True)])
        except PinError:
            continue
        if all(ins[i]["op"] == 5 for i, _, _ in f) and len(f) == len(pins) + 1:
            pins.append(([r], True))
        if len(pins) == want:
            break
    assert len(pins) == want
    return pins + [([30, 31], False)] if plan_ok(blob, pins + [([30, 31], False)]) else pins


def plan_ok(blob, pins):
    try:
        plan_pinned(blob, pins)
        return True
    except PinError:
        return False


def cases(hip):
    out = {
        "two_coins": ("H 0\nH 1\nM 0\nM 1\n", [([0, 1], True)]),
        "noisy_repeat": ("H 0\nM 0\nX_ERROR(0.25) 0\nM 0\n", [([0], True)]),
        "ghz_coins": ("H 0\nT 0\nT_DAG 0\nH 1\nCX 1 2\nH 3\nM 1 2 3\n", [([0], True), ([2], False)]),
        # a pinned active measurement: Z1 of (|00> + e^{i pi/4}|11>)/sqrt2 is fair
        "active_fair": ("H 0\nT 0\nCX 0 1\nM 1\nM 0\n", [([0], True)]),
    }
    c1 = w.surface_code_t(5, 3, 1e-3, 1, 0)
    out["surface_d5"] = (c1, surface_pins(hip.Circuit(c1).blob))
    return out


@pytest.fixture(autouse=True)
def clear_pins():
    yield
    O.set_pins(None)


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
def test_pinned_rows_match_oracle(hip, tier):
    for name, (text, pins) in cases(hip).items():
        c = hip.Circuit(text)
        s = c.compile(pin_measurements=pins, tier=tier)
        shots = 2000
        r = s.sample(shots, seed=17, bit_packed=True)
        O.set_pins(c.blob, pins)
        o = O.sample_per_shot(c.blob, shots, 17, bit_packed=True)
        print(name, tier, "pins", len(pins), flush=True)
        assert np.array_equal(r.measurements, o["measurements"]), name
        assert np.array_equal(r.detectors, o["detectors"]), name
        assert r.observable_ones.tolist() == o["observable_ones"].tolist(), name


def test_pinned_parity_holds_in_every_shot(hip):
    s = hip.Circuit("H 0\nH 1\nM 0\nM 1\n").compile(pin_measurements=[([0, 1], True)])
    m = np.unpackbits(s.sample(4096, seed=3, bit_packed=True).measurements, axis=1, bitorder="little")[:, :2]
    assert np.all(m[:, 0] ^ m[:, 1]) and 0 < m[:, 0].sum() < 4096


@pytest.mark.parametrize("tier", ["lane", "group", "wide"])
def test_a_biased_pinned_branch_fails_sampling(hip, tier):
    s = hip.Circuit("H 0\nT 0\nH 0\nM 0\n").compile(pin_measurements=[([0], True)], tier=tier)
    with pytest.raises(hip.TicitError, match="not one half"):
        s.sample(64, seed=3)


def test_decided_or_unwritten_constraints_fail_creation(hip):
    c = hip.Circuit("H 0\nM 0\nX_ERROR(0.25) 0\nM 0\n")
    with pytest.raises(hip.TicitError, match="deterministic"):
        c.compile(pin_measurements=[([0, 1], True)])
    c.compile(pin_measurements=[([0, 1], False)])
    with pytest.raises(hip.TicitError, match="not written"):
        c.compile(pin_measurements=[([9], True)])
