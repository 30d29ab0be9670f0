"""Wide (HBM-streamed) tier: parity with the oracle in FP64, and the FP32
amplitude path with FP64 reductions, norms, renormalisation and collapse
against the FP64 oracle (gpu-marked).

Tolerances: FP64 probabilities 1e-12 on these small states (and the
summation-error bound of tests/precision.py at C3's widths), FP64
amplitudes 1e-12 relative to the state norm; FP32 probabilities and
amplitudes 1e-6, unscaled.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.precision import FP32_AMP_TOL, FP32_PROB_TOL, fp64_sum_bound
from ticit_amd import workloads as w

pytestmark = pytest.mark.gpu


def random_circuit(n, layers, t, idx, noise=0.0, final_measure=True):
    rng = w.SplitMix64(w.block_seed(0xBEEF, n, idx))
    text = w.random_clifford_t(n, layers, t, rng, depolarize1=noise)
    if not final_measure:
        text = "\n".join(l for l in text.splitlines() if not l.startswith("M ")) + "\n"
    return text


CASES = {
    "c0": w.config0_circuit(),
    "k14": random_circuit(16, 20, 16, 1),
    "noisy_k9": random_circuit(12, 16, 9, 2, noise=0.02),
    "wide_compact": open(__file__.replace("test_gpu_wide.py", "golden/circuits/wide_compact.stim")).read(),
}


def away_from(probs, uni):
    close = np.abs(uni - probs) < 1e-5
    uni[close] = np.where(probs[close] < 0.5, probs[close] + 0.25, probs[close] - 0.25)
    return uni


@pytest.mark.parametrize("name", sorted(CASES))
def test_wide_fp64_seeded_rows(hip, name):
    c = hip.Circuit(CASES[name])
    s = c.compile(tier="wide", wide_batch=97)  # a batch that does not divide the shots
    assert s.tier == "wide"
    shots = 500
    r = s.sample(shots, seed=21, bit_packed=True)
    o = O.sample_per_shot(c.blob, shots, 21)
    for key in ("measurements", "detectors", "observables"):
        assert np.array_equal(getattr(r, key), o[key]), key
    assert r.observable_ones.tolist() == o["observable_ones"].tolist()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", FP32_PROB_TOL)])
@pytest.mark.parametrize("name", sorted(CASES))
def test_wide_injected_probabilities(hip, name, precision, tol):
    c = hip.Circuit(CASES[name])
    s = c.compile(tier="wide", precision=precision)
    shots = 300
    rng = np.random.default_rng(3)
    words = (shots + 63) // 64
    exo = rng.integers(0, 1 << 62, size=(c.num_symbols, words), dtype=np.uint64) & np.uint64(0x0101010101010101)
    coins = rng.integers(0, 1 << 62, size=(s.num_coin_instructions, words), dtype=np.uint64)
    uni = rng.random((s.num_measure_instructions, shots))
    for _ in range(3):
        o = O.sample_injected(c.blob, shots, exo, uni, coins, want_probs=True)
        uni = away_from(o["probs"], uni)
    o = O.sample_injected(c.blob, shots, exo, uni, coins, want_probs=True)
    r = s.sample_injected(shots, exo, uni, coins, want_probs=True)
    assert np.array_equal(r.measurements, o["measurements"])
    if o["probs"].size:
        err = np.abs(r.probabilities - o["probs"])
        print(f"{name} {precision} max probability error {float(err.max()):.3e}")
        # the FP32 path renormalizes by the stored state's norm at every
        # measurement, so storage rounding does not grow through the
        # 1 / P(branch) scaling of the collapse: the bar is unscaled
        assert float(err.max()) <= tol


def test_wide_fp32_tracks_fp32_storage_emulation(hip):
    """The FP32 path equals FP32-storage arithmetic: its probabilities agree
    with the oracle run with FP32 amplitude storage far more closely than
    either agrees with FP64."""
    c = hip.Circuit(CASES["c0"])
    s = c.compile(tier="wide", precision="fp32")
    shots = 300
    rng = np.random.default_rng(3)
    words = (shots + 63) // 64
    exo = np.zeros((c.num_symbols, words), dtype=np.uint64)
    coins = rng.integers(0, 1 << 62, size=(s.num_coin_instructions, words), dtype=np.uint64)
    uni = rng.random((s.num_measure_instructions, shots))
    o64 = O.sample_injected(c.blob, shots, exo, uni, coins, want_probs=True)
    uni = away_from(o64["probs"], uni)
    o64 = O.sample_injected(c.blob, shots, exo, uni, coins, want_probs=True)
    O.set_fp32_storage(True)
    try:
        o32 = O.sample_injected(c.blob, shots, exo, uni, coins, want_probs=True)
    finally:
        O.set_fp32_storage(False)
    r = s.sample_injected(shots, exo, uni, coins, want_probs=True)
    d_emul = float(np.max(np.abs(r.probabilities - o32["probs"])))
    d_64 = float(np.max(np.abs(r.probabilities - o64["probs"])))
    print(f"fp32 gpu vs fp32 emulation {d_emul:.3e}, vs fp64 {d_64:.3e}")
    assert d_emul <= 0.05 * d_64 + 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", FP32_AMP_TOL)])
def test_wide_final_state_amplitudes(hip, precision, tol):
    text = random_circuit(16, 14, 14, 7, final_measure=False)
    c = hip.Circuit(text)
    # the dense layout, whichever execution the planner chose
    s = c.compile(tier="wide", precision=precision, components="dense")
    O.set_component_mode(0)
    try:
        o = O.sample_injected(c.blob, 1, None, None, None, want_final_state=True)
    finally:
        O.set_component_mode(-1)
    st = s.final_state(1)
    ref = o["final_state"]
    assert st.shape == ref.shape and ref.size > 1
    err = float(np.max(np.abs(st - ref)))
    print(f"k={int(np.log2(ref.size))} {precision} max amplitude error {err:.3e}")
    assert err <= tol * float(np.linalg.norm(ref))


def c3_text(final_measure=True):
    text = w.config3_circuit()
    if not final_measure:  # drop the final M on all 30 qubits (mid-circuit M keep 8 targets)
        text = "\n".join(l for l in text.splitlines() if not (l.startswith("M ") and len(l.split()) == 31)) + "\n"
    return text


@pytest.mark.parametrize("layout,shots", [("planned", 8), ("dense", 4)])
def test_c3_precision_contract(hip, layout, shots):
    """The dense-stress configuration (30 qubits, 60 layers, 24 T, mid-circuit
    M) on its planned component layout (widest component 17) and forced
    dense (width 20): FP32 probabilities within 1e-6 of the FP64 oracle,
    FP64 within the summation-error bound, rows bit-exact."""
    c = hip.Circuit(c3_text())
    s64 = c.compile(tier="wide", components=layout)
    s32 = c.compile(tier="wide", components=layout, precision="fp32")
    rng = np.random.default_rng(3)
    uni = rng.random((s64.num_measure_instructions, shots))
    O.set_component_mode({"planned": -1, "dense": 0}[layout])
    try:
        o = O.sample_injected(c.blob, shots, None, uni, None, want_probs=True)
        uni = away_from(o["probs"], uni)
        o = O.sample_injected(c.blob, shots, None, uni, None, want_probs=True)
    finally:
        O.set_component_mode(-1)
    r64 = s64.sample_injected(shots, None, uni, None, want_probs=True)
    r32 = s32.sample_injected(shots, None, uni, None, want_probs=True)
    e64 = np.abs(r64.probabilities - o["probs"])
    e32 = np.abs(r32.probabilities - o["probs"])
    bound = fp64_sum_bound(o["probs"], uni, s64.max_local_k)
    print(f"c3 {layout} width {s64.max_local_k}: fp64 max error {e64.max():.3e} (bound {bound.min():.2e}.."
          f"{bound.max():.2e}), fp32 max error {e32.max():.3e}", flush=True)
    assert np.array_equal(r64.measurements, o["measurements"])
    assert np.array_equal(r32.measurements, o["measurements"])
    assert np.all(e64 <= bound)
    assert float(e32.max()) <= FP32_PROB_TOL


def test_c3_dense_amplitudes(hip):
    """C3 without its final M keeps a width-20 state: its FP32 amplitudes
    within 1e-6 of the FP64 oracle's, FP64 within 1e-12 relative to the
    norm."""
    c = hip.Circuit(c3_text(final_measure=False))
    O.set_component_mode(0)
    try:
        ref = O.sample_injected(c.blob, 1, None, None, None, want_final_state=True)["final_state"]
    finally:
        O.set_component_mode(-1)
    assert ref.size == 1 << 20
    for precision, tol in (("fp64", 1e-12), ("fp32", FP32_AMP_TOL)):
        st = c.compile(tier="wide", precision=precision, components="dense").final_state(1)
        err = float(np.max(np.abs(st - ref)))
        print(f"c3 width 20 {precision} max amplitude error {err:.3e}", flush=True)
        assert err <= tol * float(np.linalg.norm(ref))
