"""CPU restatement of the learned planner's MLP (test infrastructure only).

Only tests/ use this module, as the checker of the device kernels in
ticit_amd/csrc/kernels/planner_mlp.hip: same parameter layout, the same
loss (softmax cross-entropy on the dense/components choice plus
``time_weight`` times the squared error of the time output, averaged over
the minibatch) and the same Adam update (Kingma & Ba 2015, bias-corrected).
Arithmetic is float64 here; the device is float32.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 12, 64, 3
SHAPES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_H, N_OUT)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def pack(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            bound = np.sqrt(6.0 / shape[0])
            parts[name] = rng.uniform(-bound, bound, size=shape)
        else:
            parts[name] = np.zeros(shape)
    return pack(parts).astype(np.float32)


def forward(theta: np.ndarray, x: np.ndarray):
    p = unpack(np.asarray(theta, dtype=np.float64))
    a1 = x @ p["W1"] + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ p["W2"] + p["b2"]
    h2 = np.maximum(a2, 0.0)
    o = h2 @ p["W3"] + p["b3"]
    return o, (x, a1, h1, a2, h2)


def loss_and_grad(theta: np.ndarray, x: np.ndarray, y: np.ndarray, t: np.ndarray, time_weight: float):
    x = np.asarray(x, dtype=np.float64)
    o, (x, a1, h1, a2, h2) = forward(theta, x)
    B = x.shape[0]
    logits = o[:, :2]
    mx = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - mx)
    prob = e / e.sum(axis=1, keepdims=True)
    onehot = np.zeros_like(prob)
    onehot[np.arange(B), y] = 1.0
    d = o[:, 2] - t
    loss = float(np.mean(-np.log(prob[np.arange(B), y]) + time_weight * d * d))
    do = np.zeros_like(o)
    do[:, :2] = (prob - onehot) / B
    do[:, 2] = 2.0 * time_weight * d / B
    p = unpack(np.asarray(theta, dtype=np.float64))
    g = {"W3": h2.T @ do, "b3": do.sum(axis=0)}
    d2 = (do @ p["W3"].T) * (a2 > 0)
    g["W2"] = h1.T @ d2
    g["b2"] = d2.sum(axis=0)
    d1 = (d2 @ p["W2"].T) * (a1 > 0)
    g["W1"] = x.T @ d1
    g["b1"] = d1.sum(axis=0)
    return loss, pack(g)


def adam_step(theta, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mhat = m / (1 - beta1 ** step)
    vhat = v / (1 - beta2 ** step)
    return theta - lr * mhat / (np.sqrt(vhat) + eps), m, v


def predict(theta, x):
    return forward(theta, np.asarray(x, dtype=np.float64))[0]
