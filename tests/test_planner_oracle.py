"""Pins the CPU restatement of the planner MLP (oracle/mlp.py), which is the
checker of the device training and inference kernels: gradients against
central finite differences (a dropped term, wrong sign or transposed
operand shows up there), the loss against closed forms, and the Adam step
against its published first-step value. Also checks the host-side feature
extraction on programs whose features are known by construction."""
import numpy as np
import pytest

from oracle import mlp


def data(B=9, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(B, mlp.N_IN))
    y = rng.integers(0, 2, size=B)
    t = rng.normal(size=B)
    return x, y, t


def test_gradient_matches_central_differences():
    theta = mlp.init_params(3).astype(np.float64)
    x, y, t = data()
    _, g = mlp.loss_and_grad(theta, x, y, t, 0.5)
    rng = np.random.default_rng(1)
    idx = rng.choice(mlp.N_PARAMS, size=200, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        num = (mlp.loss_and_grad(tp, x, y, t, 0.5)[0] - mlp.loss_and_grad(tm, x, y, t, 0.5)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_loss_closed_form_at_zero_parameters():
    # all-zero parameters: logits 0 -> cross-entropy ln 2, time output 0
    theta = np.zeros(mlp.N_PARAMS)
    x, y, t = data(B=5)
    loss, g = mlp.loss_and_grad(theta, x, y, t, 0.25)
    assert loss == pytest.approx(np.log(2.0) + 0.25 * np.mean(t ** 2), rel=1e-12)
    b3 = mlp.unpack(g)["b3"]
    assert b3[0] == pytest.approx(np.mean(0.5 - (y == 0)), abs=1e-15)
    assert b3[2] == pytest.approx(np.mean(2 * 0.25 * (0 - t)), abs=1e-15)


def test_adam_first_step_moves_each_parameter_by_lr():
    # with m = v = 0, the bias-corrected first step is lr * g / (|g| + eps)
    theta = np.zeros(4)
    g = np.array([0.5, -2.0, 1e-3, 3.0])
    new, m, v = mlp.adam_step(theta, np.zeros(4), np.zeros(4), g, 1, lr=0.01)
    assert np.allclose(new, -0.01 * np.sign(g), rtol=1e-4)


def test_parameter_layout():
    assert mlp.N_PARAMS == 5187
    parts = mlp.unpack(np.arange(mlp.N_PARAMS, dtype=np.float64))
    assert parts["W1"][0, 1] == 1 and parts["b1"][0] == 12 * 64
    assert parts["W3"][0, 2] == 12 * 64 + 64 + 64 * 64 + 64 + 2


def test_features_of_known_programs():
    import ticit_amd
    from tests import blobs as B
    # separable 12-qubit program: 40 diagonal rotations, no promotions
    k = 12
    ins = [B.rotation(k, zmask=1 << (s % k), angle=0.05 * (s + 1)) for s in range(40)]
    f = ticit_amd.planner_features(B.encode(k, k, k, 0, 0, 0, ins))
    assert f["qubits"] == 12 and f["t_count"] == 40 and f["mean_log2_active"] == 12
    assert f["nonzero_fraction"] == pytest.approx(24 / 4096)   # 12 live components of 2
    assert f["log2_dense_peak"] == 12 and f["quantum_ops"] == 40
    assert f["log2_dense_work"] == pytest.approx(np.log2(1 + 40 * 4096))
    assert f["log2_component_work"] == pytest.approx(np.log2(1 + 40 * 34))
    # a circuit: layers and records come from the text
    text = "H 0\nTICK\nT 0\nTICK\nCX 0 1\nM 0 1\n"
    g = ticit_amd.planner_features(ticit_amd.plan_text(text))
    assert g["clifford_depth"] == 3 and g["measurements"] == 2 and g["qubits"] == 2
