"""Pinned measurement parities in the CPU oracle (oracle/pinning.py and the
forced-branch execution of oracle/ref_cpu.cpp), pinned by the reference's
own tests: the solver cases of src/sampler/pinning.rs:300-387 and the
sampling behaviour of src/sampler/prepared.rs:1389-1530."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle.pinning import PinError, plan_pinned
from ticit_amd import plan_text


@pytest.fixture(autouse=True)
def clear_pins():
    yield
    O.set_pins(None)


def test_a_free_coin_is_pinned_at_its_own_instruction():
    forced = plan_pinned(plan_text("H 0\nM 0\n"), [([0], True)])
    assert len(forced) == 1 and forced[0][1] is True and forced[0][2] == []


def test_a_deterministic_record_rejects_the_wrong_value_and_accepts_the_right_one():
    with pytest.raises(PinError, match="deterministic"):
        plan_pinned(plan_text("M 0\n"), [([0], True)])
    assert plan_pinned(plan_text("M 0\n"), [([0], False)]) == []


def test_a_parity_pins_only_its_last_free_coin():
    forced = plan_pinned(plan_text("H 0\nH 1\nM 0\nM 1\n"), [([0, 1], True)])
    assert len(forced) == 1 and len(forced[0][2]) == 1


def test_independent_constraints_take_distinct_pivots():
    forced = plan_pinned(plan_text("H 0\nH 1\nM 0\nM 1\n"), [([0], True), ([0, 1], False)])
    assert len(forced) == 2 and forced[0][0] != forced[1][0]


def test_contradictory_constraints_and_unwritten_records_are_rejected():
    with pytest.raises(PinError, match="deterministic"):
        plan_pinned(plan_text("H 0\nM 0\nM 0\n"), [([0], True), ([1], False)])
    with pytest.raises(PinError, match="not written"):
        plan_pinned(plan_text("H 0\nM 0\n"), [([7], True)])


def bits(rows, ncols):
    return np.unpackbits(rows, axis=1, bitorder="little")[:, :ncols]


@pytest.mark.parametrize("value", [False, True])
def test_pinning_holds_the_parity_in_every_shot(value):
    blob = plan_text("H 0\nH 1\nM 0\nM 1\n")
    O.set_pins(blob, [([0, 1], value)])
    m = bits(O.sample(blob, 512, 9)["measurements"], 2)
    assert np.all((m[:, 0] ^ m[:, 1]) == int(value))
    assert 0 < m[:, 0].sum() < 512  # only the parity is pinned


def test_pinning_leaves_noise_on_the_records():
    blob = plan_text("H 0\nM 0\nX_ERROR(0.25) 0\nM 0\n")
    O.set_pins(blob, [([0], True)])
    m = bits(O.sample(blob, 4096, 11)["measurements"], 2)
    assert np.all(m[:, 0] == 1)
    assert abs((m[:, 1] != 1).mean() - 0.25) < 0.03


def test_a_deterministic_parity_only_accepts_its_own_value():
    blob = plan_text("H 0\nM 0\nX_ERROR(0.25) 0\nM 0\n")
    O.set_pins(blob, [([0, 1], False)])
    with pytest.raises(PinError, match="deterministic"):
        O.set_pins(blob, [([0, 1], True)])


def test_pinning_a_biased_branch_is_refused():
    blob = plan_text("H 0\nT 0\nH 0\nM 0\n")
    O.set_pins(blob, [([0], True)])
    with pytest.raises(O.OracleError, match="not one half"):
        O.sample(blob, 8, 3)


def test_pinning_both_values_reproduces_the_unpinned_distribution():
    text = "H 0\nCX 0 1\nM 0\nX_ERROR(0.25) 1\nM 1\nOBSERVABLE_INCLUDE(0) rec[-1]\n"
    blob = plan_text(text)
    shots = 100_000
    free = O.sample(blob, 2 * shots, 5, threads=8)["counts"]
    free_rate = int(free[3]) / (2 * shots)
    ones = 0
    for value in (False, True):
        O.set_pins(blob, [([0], value)])
        ones += int(O.sample(blob, shots, 5, threads=8)["counts"][3])
    assert abs(ones / (2 * shots) - free_rate) < 0.008
