"""Tensor-component planning and execution in the oracle.

The programs are the reference's own hand-built ones
(src/component_plan.rs:560-734); the asserted values are the ones its tests
assert. The oracle's component execution is then checked against its dense
execution on the same draws, which is what src/component_plan.rs:689-700
asserts for its program (identical records at seed 43).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import blobs as B


@pytest.fixture(autouse=True)
def _dense_by_default():
    O.set_component_mode(-1)
    yield
    O.set_component_mode(-1)


def component_test_program(use_components: bool = False) -> np.ndarray:
    # src/component_plan.rs:576-611: n = 8, initial_k = 0, max_k = 8
    ins = [
        B.rotation(0, angle=0.19),
        B.promotion(0, 0.17),
        B.promotion(1, -0.29),
        B.promotion(2, 0.41),
        B.rotation(3, xmask=0b110, angle=0.23),    # merge_late: X1 X2
        B.rotation(3, xmask=0b011, angle=-0.31),   # merge_early_into_late: X0 X1
        B.measurement(3, xmask=0b011, branch=1, record=1),
        B.detector(1, [1]),
        B.measurement(2, zmask=0b10, branch=2, record=2),  # reversed_single: Z1
        B.measurement(1, xmask=0b1, branch=3, record=3),   # final_x: X0
    ]
    return B.encode(8, 0, 8, 3, 3, 1, ins, use_components=use_components)


def separable_program(use_components: bool = True) -> np.ndarray:
    # src/component_plan.rs:703-724: k = 12, 40 single-qubit Z rotations
    k = 12
    ins = [B.rotation(k, zmask=1 << (step % k), angle=0.05 * (step + 1)) for step in range(40)]
    return B.encode(k, k, k, 0, 0, 0, ins, use_components=use_components)


def test_small_active_states_are_never_factored():
    # src/component_plan.rs:612-619
    plan = O.component_plan(B.encode(4, 4, 4, 0, 0, 0, []))
    assert not plan["selected"]
    assert not plan["planned"] and plan["steps"] == []
    assert plan["dense_peak_dimension"] == 16


def test_component_planner_tracks_merges_and_local_pivots():
    # src/component_plan.rs:647-700
    plan = O.component_plan(component_test_program())
    kinds = [s["kind"] for s in plan["steps"]]
    assert plan["initial_components"] == 0
    assert kinds.count("promotion") == 3
    assert kinds.count("rotation") == 2
    assert kinds.count("measurement") == 3
    assert kinds[0] == "ignored_global_phase"
    assert kinds[7] == "none"
    rots = [s for s in plan["steps"] if s["kind"] == "rotation"]
    meas = [s for s in plan["steps"] if s["kind"] == "measurement"]
    m = plan["merges"]
    assert rots[0]["component"] == 1
    assert m[rots[0]["merge_offset"]:][:rots[0]["merge_count"]] == [1, 2]
    assert rots[1]["component"] == 1
    assert m[rots[1]["merge_offset"]:][:rots[1]["merge_count"]] == [1, 0]
    assert meas[0]["component"] == 1
    assert m[meas[0]["merge_offset"]:][:meas[0]["merge_count"]] == [1]
    assert meas[0]["pivot"] == 0
    assert meas[2]["deactivate_after"]
    assert plan["component_max_k"] == [1, 3, 1]
    assert not plan["selected"]


def test_component_records_equal_dense_records_at_seed_43():
    # src/component_plan.rs:689-700: 256 shots, seed 43
    dense = O.sample(component_test_program(False), 256, 43, bit_packed=False)
    comp = O.sample(component_test_program(True), 256, 43, bit_packed=False)
    assert np.array_equal(dense["measurements"], comp["measurements"])
    assert np.array_equal(dense["detectors"], comp["detectors"])
    # a physical sanity check on top: the X0X1 outcome is not constant
    assert 0 < dense["measurements"][:, 0].sum() < 256


def test_fully_separable_program_selects_the_component_plan():
    # src/component_plan.rs:703-724
    plan = O.component_plan(separable_program())
    assert plan["selected"]
    assert plan["initial_components"] == 12
    assert plan["component_max_k"] == [1] * 12
    assert plan["dense_peak_dimension"] == 1 << 12
    assert plan["component_allocated_dimension"] == 24
    assert plan["estimated_dense_work"] > plan["estimated_component_work"]
    # closed forms of the cost model for this program: 40 rotations at
    # 2^12 dense; 40 x (2^1 + 32) component, no merges
    assert plan["estimated_dense_work"] == 40 * 4096
    assert plan["estimated_component_work"] == 40 * (2 + 32)


def test_expectation_probes_disable_the_component_plan():
    # src/component_plan.rs:726-733
    ins = [B.rotation(0, angle=0.19), B.promotion(0, 0.17)]
    plan = O.component_plan(B.encode(8, 0, 8, 0, 0, 0, ins, nexpvals=1))
    assert not plan["planned"] and not plan["selected"]


def test_merge_work_and_peaks_follow_the_cost_model():
    # by hand for component_test_program: dense work 2^0 (rotation at k=0)
    # + 2(1+2+4) promotions + 8 + 8 rotations + 2(8+4+2) measurements;
    # component work: merges 2*4 and 2*8, ops 4 + 8, promotions 3*2,
    # measurements 2(8+4+2), dispatch 32 per non-ignored quantum op
    plan = O.component_plan(component_test_program())
    assert plan["estimated_dense_work"] == 1 + 2 * (1 + 2 + 4) + 8 + 8 + 2 * (8 + 4 + 2)
    assert plan["estimated_component_work"] == (2 * 4 + 2 * 8) + (4 + 8) + 3 * 2 + 2 * (8 + 4 + 2) + 32 * 8
    assert plan["component_allocated_dimension"] == 2 + 8 + 2
    # live sizes: 3 x 2 after the promotions, 4 + 2 after the first merge,
    # 8 once all three coordinates share one component
    assert plan["component_peak_live_dimension"] == 8
    assert plan["dense_peak_dimension"] == 256


def _injected(blob, shots, seed=5):
    inf = O.info(blob)
    rng = np.random.default_rng(seed)
    words = (shots + 63) // 64
    uni = rng.random((max(1, inf["nmeasure"]), shots))
    coins = rng.integers(0, 1 << 62, size=(max(1, inf["ncoin"]), words), dtype=np.uint64)
    return O.sample_injected(blob, shots, None, uni, coins, want_probs=True, bit_packed=False)


@pytest.mark.parametrize("which", ["test_program", "separable"])
def test_component_probabilities_match_dense(which):
    blob = component_test_program(True) if which == "test_program" else separable_program(True)
    O.set_component_mode(0)
    dense = _injected(blob, 64)
    O.set_component_mode(1)
    comp = _injected(blob, 64)
    assert np.array_equal(dense["measurements"], comp["measurements"])
    if "probs" in dense and dense["probs"].size:
        assert np.max(np.abs(dense["probs"] - comp["probs"])) <= 1e-12


def test_compensated_sums_agree_with_the_scalar_order_on_small_states():
    # a k <= 12 program: both orders are within a few ulps of each other
    from ticit_amd import plan_text
    from ticit_amd import workloads as W
    blob = plan_text(W.config0_circuit())
    rng = np.random.default_rng(4)
    inf = O.info(blob)
    uni = rng.random((inf["nmeasure"], 16))
    a = O.sample_injected(blob, 16, None, uni, None, want_probs=True)
    O.set_exact_sums(True)
    try:
        b = O.sample_injected(blob, 16, None, uni, None, want_probs=True)
    finally:
        O.set_exact_sums(False)
    assert np.array_equal(a["measurements"], b["measurements"])
    assert np.max(np.abs(a["probs"] - b["probs"])) <= 1e-13
