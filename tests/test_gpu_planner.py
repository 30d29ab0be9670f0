"""Learned planner kernels on the GPU (gpu-marked), checked against the CPU
restatement in oracle/mlp.py: gradients of the mean minibatch loss, one
Adam step, inference, a short training run on a learnable rule, and the
layout decision at program creation with its heuristic fallback.

Tolerances: FP32 device arithmetic against FP64 on the CPU; gradients
within 1e-4 of the largest gradient entry, outputs within 1e-4 absolute.
"""
import numpy as np
import pytest

from oracle import mlp
from ticit_amd import planner as PL
from ticit_amd import workloads as W

pytestmark = pytest.mark.gpu


def batch(B, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(B, 12)).astype(np.float32)
    y = rng.integers(0, 2, size=B).astype(np.int32)
    t = rng.normal(size=B).astype(np.float32)
    return x, y, t


@pytest.mark.parametrize("B", [1, 37, 64])
def test_gradients_match_the_oracle(hip, B):
    theta = PL.init_params(5)
    net = PL.DeviceMLP(theta)
    x, y, t = batch(B, B)
    loss, g = net.gradients(x, y, t, 0.5)
    ref_loss, ref_g = mlp.loss_and_grad(theta.astype(np.float64), x.astype(np.float64), y, t.astype(np.float64), 0.5)
    err = float(np.max(np.abs(g - ref_g)))
    print(f"B={B}: loss {loss:.6f} vs {ref_loss:.6f}, max gradient error {err:.2e} of {np.max(np.abs(ref_g)):.2e}")
    assert abs(loss - ref_loss) <= 1e-5 * max(1.0, abs(ref_loss))
    assert err <= 1e-4 * float(np.max(np.abs(ref_g)))
    assert np.array_equal(net.params(), theta)  # gradients only, no update


def test_one_adam_step_matches_the_oracle(hip):
    theta = PL.init_params(9)
    net = PL.DeviceMLP(theta)
    x, y, t = batch(64, 3)
    net.train(x, y, t, np.arange(64, dtype=np.uint32), 64, 1e-3, 0.5)
    _, g = mlp.loss_and_grad(theta.astype(np.float64), x.astype(np.float64), y, t.astype(np.float64), 0.5)
    ref, _, _ = mlp.adam_step(theta.astype(np.float64), np.zeros(mlp.N_PARAMS), np.zeros(mlp.N_PARAMS), g, 1,
                              lr=1e-3)
    got = net.params()
    # first Adam step: lr * g / (|g| + eps) -- insensitive to gradient
    # scale, so compare only where |g| is not at the eps level
    mask = np.abs(g) > 1e-6
    assert np.max(np.abs(got[mask] - ref[mask])) <= 1e-6


def test_inference_matches_the_oracle(hip):
    theta = PL.init_params(11)
    net = PL.DeviceMLP(theta)
    x, _, _ = batch(300, 4)
    out = net.predict(x)
    ref = mlp.predict(theta, x)
    assert np.max(np.abs(out - ref)) <= 1e-4


def test_training_learns_a_rule(hip):
    rng = np.random.default_rng(2)
    x = rng.normal(size=(4096, 12)).astype(np.float32)
    y = (x[:, 0] + 0.5 * x[:, 3] > 0).astype(np.int32)
    t = (0.7 * x[:, 1] - 0.2 * x[:, 2]).astype(np.float32)
    net = PL.DeviceMLP(PL.init_params(1))
    first = last = None
    for ep in range(20):
        losses = net.train(x, y, t, rng.permutation(4096).astype(np.uint32), 64, 2e-3, 0.5)
        first = losses[:8].mean() if first is None else first
        last = losses[-8:].mean()
    out = net.predict(x)
    acc = float(np.mean((out[:, 1] > out[:, 0]) == (y == 1)))
    print(f"loss {first:.3f} -> {last:.3f}, accuracy {acc:.3f}")
    assert last < 0.25 * first and acc > 0.95


def test_learned_decision_and_fallback(hip):
    from ticit_amd import Circuit
    blob = Circuit(W.config2_batch()[235][0]).blob  # planned, heuristic selects components
    theta = PL.init_params(3)
    feats = PL.feature_vector(blob)
    mean, scale = np.zeros(12, np.float32), np.ones(12, np.float32)
    lib = hip._capi.load()
    fp = lambda a: a.ctypes.data_as(PL._f32p)
    try:
        hip._capi.check(lib.ticit_hip_planner_set(fp(theta), fp(mean), fp(scale), 0.0))
        s = hip.CompiledSampler(blob, planner="heuristic")
        assert s.planner_source == "heuristic" and s.components
        # planner_mode 1 with the test model: the decision follows its output
        from ticit_amd import _capi
        import ctypes
        opt = _capi.Options()
        opt.planner_mode = 1
        h = ctypes.c_void_p()
        _capi.check(lib.ticit_hip_program_create(blob.ctypes.data_as(_capi._u64p), blob.size, ctypes.byref(opt),
                                                 ctypes.byref(h)))
        info = _capi.ProgramInfo()
        _capi.check(lib.ticit_hip_program_info_get(h, ctypes.byref(info)))
        lib.ticit_hip_program_destroy(h)
        ref = mlp.predict(theta, feats[None, :])[0]
        p = float(np.exp(ref[1]) / (np.exp(ref[0]) + np.exp(ref[1])))
        assert info.planner_source == 1
        assert abs(info.planner_p_components - p) <= 1e-4
        assert bool(info.components) == (p > 0.5)
        # a margin no confidence reaches: the heuristic's choice stands
        hip._capi.check(lib.ticit_hip_planner_set(fp(theta), fp(mean), fp(scale), 0.6))
        h = ctypes.c_void_p()
        _capi.check(lib.ticit_hip_program_create(blob.ctypes.data_as(_capi._u64p), blob.size, ctypes.byref(opt),
                                                 ctypes.byref(h)))
        _capi.check(lib.ticit_hip_program_info_get(h, ctypes.byref(info)))
        lib.ticit_hip_program_destroy(h)
        assert info.planner_source == 0 and bool(info.components) == bool(int(blob[9]) & 1)
    finally:
        PL.uninstall()


def test_held_out_batch_runs_the_shipped_model_on_the_device(hip):
    """The C4 workload: the 512 held-out programs created with
    planner="learned". The device inference's P(components) equals the
    oracle's forward pass of the shipped model (1e-4), the layout each
    program runs on is the decision rule's (learned choice, or the
    heuristic's when |p - 1/2| is under the margin), and the number of
    fallbacks is the stored evaluation's; a sample of the programs then
    samples the same rows batched as one at a time."""
    import torch

    from oracle import mlp
    from ticit_amd import planner as PL
    from ticit_amd import workloads as W

    m = PL.load_model()
    x = (m["ho_x"] - m["mean"]) / m["scale"]
    p_ref = PL.softmax_p_components(mlp.predict(m["params"], x))
    want, confident = PL.decide(p_ref, m["ho_heuristic"], float(m["margin"]))
    samplers = []
    for j, idx in enumerate(m["ho_idx"]):
        s = hip.CompiledSampler(hip.plan_text(W.planner_corpus_circuit(int(idx))), planner="learned")
        samplers.append(s)
        assert abs(s.planner_p_components - float(p_ref[j])) <= 1e-4, j
        assert s.components == bool(want[j]), j
        assert (s.planner_source == "learned") == bool(confident[j]), j
    fallbacks = sum(s.planner_source != "learned" for s in samplers)
    print("held-out programs", len(samplers), "components", sum(s.components for s in samplers),
          "fallbacks", fallbacks, flush=True)
    assert fallbacks == int((~confident).sum())
    pick = samplers[::32]
    shots = [3000] * len(pick)
    row = lambda n: max(1, (n + 7) // 8)
    outs, singles = [], []
    for s, n in zip(pick, shots):
        mk = lambda: (torch.zeros((n, row(s.num_measurements)), dtype=torch.uint8, device="cuda"),
                      torch.zeros((n, row(s.num_detectors)), dtype=torch.uint8, device="cuda"),
                      torch.zeros((n, row(s.num_observables)), dtype=torch.uint8, device="cuda"),
                      torch.zeros(max(1, s.num_observables), dtype=torch.int64, device="cuda"))
        outs.append(mk())
        singles.append(mk())
    b = hip.Batch(pick, shots, list(range(len(pick))), outs, autotune=True, autotune_shots=512)
    b.launch(0)
    b.wait()
    for i, (s, n) in enumerate(zip(pick, shots)):
        s.sample_device(n, i, singles[i])
    torch.cuda.synchronize()
    for a, c in zip(outs, singles):
        assert all(torch.equal(u, v) for u, v in zip(a, c))
    b.close()
