"""Tensor-component execution on the GPU (gpu-marked).

The reference runs a program on its component plan whenever the planner's
cost model selects it (src/planner.rs:1482-1484,
src/sampler/batch/runtime.rs:1094-1098), so a selected program's rows are
the component path's rows. Checked here against the oracle's component
execution: rows bit-exact on seeded and injected draws, P(branch = 1) within
1e-12; and forced-dense against forced-component runs on the GPU within
1e-12 (the two layouts compute the same state in a different order).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import blobs as B
from tests.precision import FP32_PROB_TOL, fp64_sum_bound
from ticit_amd import workloads as W

pytestmark = pytest.mark.gpu


def _c2(i):
    return W.config2_batch()[i][0]


def separable_blob():
    k = 12
    ins = [B.rotation(k, zmask=1 << (step % k), angle=0.05 * (step + 1)) for step in range(40)]
    # X on the top active qubit, k shrinking by one each time
    ins += [B.measurement(q + 1, xmask=1 << q, branch=q + 1, record=q + 1) for q in reversed(range(k))]
    return B.encode(k, k, k, k, k, 0, ins, use_components=True)


def component_test_blob():
    # src/component_plan.rs:576-611
    ins = [
        B.rotation(0, angle=0.19),
        B.promotion(0, 0.17),
        B.promotion(1, -0.29),
        B.promotion(2, 0.41),
        B.rotation(3, xmask=0b110, angle=0.23),
        B.rotation(3, xmask=0b011, angle=-0.31),
        B.measurement(3, xmask=0b011, branch=1, record=1),
        B.detector(1, [1]),
        B.measurement(2, zmask=0b10, branch=2, record=2),
        B.measurement(1, xmask=0b1, branch=3, record=3),
    ]
    return B.encode(8, 0, 8, 3, 3, 1, ins, use_components=True)


TEXT_CASES = {"c2_094": 94, "c2_214": 214, "c2_235": 235}


def case_blob(hip, name):
    if name == "separable":
        return separable_blob()
    if name == "test_program":
        return component_test_blob()
    if name == "c3":
        return hip.Circuit(W.config3_circuit()).blob
    return hip.Circuit(_c2(TEXT_CASES[name])).blob


@pytest.fixture(autouse=True)
def _oracle_follows_blob():
    O.set_component_mode(-1)
    yield
    O.set_component_mode(-1)


def away_from(probs, uni):
    close = np.abs(uni - probs) < 1e-6
    uni[close] = np.where(probs[close] < 0.5, probs[close] + 0.25, probs[close] - 0.25)
    return uni


@pytest.mark.parametrize("name,tier", [("test_program", "group"), ("test_program", "wide"),
                                       ("separable", "group"), ("separable", "wide"),
                                       ("c2_094", "group"), ("c2_214", "group"), ("c2_214", "wide"),
                                       ("c2_235", "wide")])
def test_component_seeded_rows_match_oracle(hip, name, tier):
    blob = case_blob(hip, name)
    assert int(blob[9]) & 1, "the planner selected the component execution"
    s = hip.CompiledSampler(blob, tier=tier, wide_batch=61)
    assert s.components and s.tier == tier
    shots = 400 if name != "c2_235" else 120
    print(name, tier, "state_dim", s.state_dim, "components", s.component_count, "max local k", s.max_local_k,
          flush=True)
    r = s.sample(shots, seed=17, bit_packed=True)
    o = O.sample_per_shot(blob, shots, 17)
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(r, key), o[key]), key


@pytest.mark.parametrize("name,tier,shots", [("c2_214", "group", 200), ("c2_235", "wide", 64), ("c3", "wide", 6)])
def test_component_injected_probabilities(hip, name, tier, shots):
    blob = case_blob(hip, name)
    s = hip.CompiledSampler(blob, tier=tier)
    assert s.components
    rng = np.random.default_rng(11)
    words = (shots + 63) // 64
    exo = rng.integers(0, 1 << 62, size=(max(1, s.num_symbols), words), dtype=np.uint64) & np.uint64(
        0x0001000100010001)
    coins = rng.integers(0, 1 << 62, size=(max(1, s.num_coin_instructions), words), dtype=np.uint64)
    uni = rng.random((max(1, s.num_measure_instructions), shots))
    for _ in range(2):
        o = O.sample_injected(blob, shots, exo, uni, coins, want_probs=True)
        uni = away_from(o["probs"], uni)
    o = O.sample_injected(blob, shots, exo, uni, coins, want_probs=True)
    r = s.sample_injected(shots, exo, uni, coins, want_probs=True)
    assert np.array_equal(r.measurements, o["measurements"])
    err = float(np.max(np.abs(r.probabilities - o["probs"])))
    print(f"{name} {tier} components: max probability error {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("name", ["c2_214", "c2_235"])
def test_dense_and_component_layouts_agree(hip, name):
    blob = case_blob(hip, name)
    shots = 96
    rng = np.random.default_rng(5)
    words = (shots + 63) // 64
    dense = hip.CompiledSampler(blob, components="dense")
    comp = hip.CompiledSampler(blob, components="components")
    assert not dense.components and comp.components
    assert comp.state_dim < dense.state_dim
    uni = rng.random((dense.num_measure_instructions, shots))
    exo = np.zeros((max(1, dense.num_symbols), words), dtype=np.uint64)
    coins = rng.integers(0, 1 << 62, size=(max(1, dense.num_coin_instructions), words), dtype=np.uint64)
    a = dense.sample_injected(shots, exo, uni, coins, want_probs=True)
    uni = away_from(a.probabilities, uni)
    a = dense.sample_injected(shots, exo, uni, coins, want_probs=True)
    b = comp.sample_injected(shots, exo, uni, coins, want_probs=True)
    assert np.array_equal(a.measurements, b.measurements)
    assert float(np.max(np.abs(a.probabilities - b.probabilities))) <= 1e-12


def test_component_fp32_path(hip):
    blob = case_blob(hip, "c2_235")
    s = hip.CompiledSampler(blob, precision="fp32")
    assert s.components and s.tier == "wide"
    shots = 32
    rng = np.random.default_rng(9)
    words = (shots + 63) // 64
    uni = rng.random((s.num_measure_instructions, shots))
    exo = np.zeros((max(1, s.num_symbols), words), dtype=np.uint64)
    coins = rng.integers(0, 1 << 62, size=(max(1, s.num_coin_instructions), words), dtype=np.uint64)
    o = O.sample_injected(blob, shots, exo, uni, coins, want_probs=True)
    uni = away_from(o["probs"], uni)
    o = O.sample_injected(blob, shots, exo, uni, coins, want_probs=True)
    O.set_fp32_storage(True)
    try:
        o32 = O.sample_injected(blob, shots, exo, uni, coins, want_probs=True)
    finally:
        O.set_fp32_storage(False)
    r = s.sample_injected(shots, exo, uni, coins, want_probs=True)
    d64 = float(np.max(np.abs(r.probabilities - o["probs"])))
    e32 = float(np.max(np.abs(o32["probs"] - o["probs"])))
    d32 = float(np.max(np.abs(r.probabilities - o32["probs"])))
    print(f"components fp32: vs fp64 {d64:.3e} (storage emulation vs fp64 {e32:.3e}), "
          f"vs emulation {d32:.3e}", flush=True)
    # FP32 storage with every measurement renormalized by the stored norm:
    # within the 1e-6 target of the FP64 oracle, rows equal
    assert np.array_equal(r.measurements, o["measurements"])
    assert d64 <= FP32_PROB_TOL


def test_planned_layout_is_the_reference_choice(hip):
    # selected programs run on components, unselected ones dense
    for i in (94, 235, 0, 1):
        c = hip.Circuit(_c2(i))
        s = c.compile()
        assert s.components == bool(int(c.blob[9]) & 1) == O.component_plan(c.blob)["selected"]


def test_fp64_probabilities_against_compensated_sums(hip):
    """C3 at width 17: the device's tree-reduced FP64 probabilities against
    the oracle with Neumaier-compensated sums (closer to exact than the
    reference's scalar order); also reports the scalar order's own distance."""
    blob = case_blob(hip, "c3")
    s = hip.CompiledSampler(blob)
    shots = 8
    rng = np.random.default_rng(3)
    uni = rng.random((s.num_measure_instructions, shots))
    plain = O.sample_injected(blob, shots, None, uni, None, want_probs=True)
    O.set_exact_sums(True)
    try:
        exact = O.sample_injected(blob, shots, None, uni, None, want_probs=True)
    finally:
        O.set_exact_sums(False)
    r = s.sample_injected(shots, None, uni, None, want_probs=True)
    assert np.array_equal(r.measurements, exact["measurements"])
    d_gpu = float(np.max(np.abs(r.probabilities - exact["probs"])))
    d_scalar = float(np.max(np.abs(plain["probs"] - exact["probs"])))
    print(f"c3 fp64: device vs compensated {d_gpu:.2e}; scalar order vs compensated {d_scalar:.2e}", flush=True)
    assert d_gpu <= 1e-12
    # against the reference's own scalar order: within the summation-error
    # bound of the two orders (tests/precision.py)
    assert np.array_equal(r.measurements, plain["measurements"])
    bound = fp64_sum_bound(plain["probs"], uni, s.max_local_k)
    assert np.all(np.abs(r.probabilities - plain["probs"]) <= bound)
