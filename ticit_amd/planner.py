"""Learned planner: which active-state layout (dense or tensor components)
runs a program faster on this device.

The decision it learns is the one ticit's planner makes with a fixed cost
model, should_select_component_plan (src/component_plan.rs:282-298), whose
choice every plan blob carries (use_active_components). The learned model
is a 12 -> 64 -> 64 -> 3 FP32 MLP trained on the device by the library's
own kernels (ticit_hip_mlp_*) on labels measured by timing both layouts;
at program creation (planner="learned") it runs as a device inference
kernel and falls back to the heuristic's choice when it is not confident.

``python -m ticit_amd.planner train`` builds the corpus, measures the
labels on the current device, trains, evaluates regret on the held-out
batch and writes the model file shipped with the package.
"""
from __future__ import annotations

import ctypes
import os
import sys
import time

import numpy as np

from . import _capi

FEATURE_NAMES = [
    "qubits", "t_count", "mean_log2_active", "nonzero_fraction", "clifford_depth", "measurements",
    "log2_dense_work", "log2_component_work", "log2_dense_peak", "log2_peak_live", "log2_allocated",
    "quantum_ops",
]
N_FEATURES = len(FEATURE_NAMES)
N_PARAMS = 5187
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planner_model.npz")

_f32p = ctypes.POINTER(ctypes.c_float)


def _fp(a):
    return a.ctypes.data_as(_f32p)


def feature_vector(blob: np.ndarray) -> np.ndarray:
    """Raw planner features of a plan blob (host only)."""
    lib = _capi.load()
    blob = np.ascontiguousarray(blob, dtype=np.uint64)
    out = np.zeros(N_FEATURES, dtype=np.float32)
    _capi.check(lib.ticit_hip_planner_features(blob.ctypes.data_as(_capi._u64p), blob.size, _fp(out)))
    return out


def planner_features(blob: np.ndarray) -> dict:
    return dict(zip(FEATURE_NAMES, (float(v) for v in feature_vector(blob))))


class DeviceMLP:
    """The MLP's parameters on the device with its training, gradient and
    inference kernels."""

    def __init__(self, params: np.ndarray):
        self._lib = _capi.load()
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters")
        h = ctypes.c_void_p()
        _capi.check(self._lib.ticit_hip_mlp_create(_fp(p), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ticit_hip_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, dtype=np.float32)
        _capi.check(self._lib.ticit_hip_mlp_params_get(self._h, _fp(out)))
        return out

    def train(self, x, y, t, order, batch: int, lr: float, time_weight: float) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        order = np.ascontiguousarray(order, dtype=np.uint32)
        steps = order.size // batch
        losses = np.zeros(max(1, steps), dtype=np.float32)
        _capi.check(self._lib.ticit_hip_mlp_train(
            self._h, _fp(x), y.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _fp(t), x.shape[0],
            order.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), steps, batch, lr, time_weight, _fp(losses)))
        return losses[:steps]

    def gradients(self, x, y, t, time_weight: float):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(N_PARAMS, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _capi.check(self._lib.ticit_hip_mlp_gradients(
            self._h, _fp(x), y.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _fp(t), x.shape[0], time_weight,
            _fp(g), _fp(loss)))
        return float(loss[0]), g

    def predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((x.shape[0], 3), dtype=np.float32)
        _capi.check(self._lib.ticit_hip_mlp_predict(self._h, _fp(x), x.shape[0], _fp(out)))
        return out


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases, in the device layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, shape in ((12, (12, 64)), (0, (64,)), (64, (64, 64)), (0, (64,)), (64, (64, 3)), (0, (3,))):
        if fan_in:
            b = np.sqrt(6.0 / fan_in)
            parts.append(rng.uniform(-b, b, size=shape).ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    return np.concatenate(parts).astype(np.float32)


def softmax_p_components(out: np.ndarray) -> np.ndarray:
    lo = out[:, :2] - out[:, :2].max(axis=1, keepdims=True)
    e = np.exp(lo)
    return e[:, 1] / e.sum(axis=1)


def decide(p_components: np.ndarray, heuristic: np.ndarray, margin: float):
    """The device decision rule: learned choice unless |p - 1/2| < margin."""
    confident = np.abs(p_components - 0.5) >= margin
    return np.where(confident, p_components > 0.5, heuristic), confident


def load_model(path: str = MODEL_PATH) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def install(path: str = MODEL_PATH) -> dict:
    """Loads the shipped model into the library (planner_mode = 1)."""
    m = load_model(path)
    lib = _capi.load()
    _capi.check(lib.ticit_hip_planner_set(_fp(np.ascontiguousarray(m["params"], np.float32)),
                                          _fp(np.ascontiguousarray(m["mean"], np.float32)),
                                          _fp(np.ascontiguousarray(m["scale"], np.float32)),
                                          float(m["margin"])))
    return m


_installed = False


def ensure_installed() -> None:
    """Installs the shipped model once per process (planner="learned")."""
    global _installed
    if not _installed:
        if not os.path.exists(MODEL_PATH):
            raise FileNotFoundError(f"no trained planner model at {MODEL_PATH}")
        install(MODEL_PATH)
        _installed = True


def uninstall() -> None:
    global _installed
    _installed = False
    _capi.check(_capi.load().ticit_hip_planner_set(None, None, None, 0.0))


# --------------------------------------------------------------------------
# Corpus, labels, training
# --------------------------------------------------------------------------

def _plan_one(index: int):
    from . import plan_text
    from .workloads import planner_corpus_circuit
    blob = plan_text(planner_corpus_circuit(index))
    f = feature_vector(blob)
    planned = f[8] >= 8  # log2 dense peak = max_k: components need max_k >= 8
    return index, (blob if planned else None), f


def planned_corpus(start: int, count: int, workers: int = 8):
    """The first `count` programs from corpus index `start` on that have a
    component plan (max_k >= 8), as (indices, blobs, raw features)."""
    from multiprocessing import get_context
    idx, blobs, feats = [], [], []
    nxt = start
    with get_context("fork").Pool(workers) as pool:
        while len(idx) < count:
            chunk = range(nxt, nxt + max(64, 2 * (count - len(idx))))
            nxt = chunk.stop
            for i, b, f in pool.imap(_plan_one, chunk, chunksize=16):
                if b is not None and len(idx) < count:
                    idx.append(i)
                    blobs.append(b)
                    feats.append(f)
    return np.array(idx), blobs, np.array(feats, dtype=np.float32)


def time_per_shot(blob: np.ndarray, components: str, seed: int = 1) -> float:
    """Device time per shot of the program in one layout, on a shot count
    sized to its state (sampling kernels only, HIP events)."""
    from . import CompiledSampler
    s = CompiledSampler(blob, components=components)
    shots = int(np.clip((1 << 24) // max(1, s.state_dim), 16, 2048))
    s.sample(max(8, shots // 8), seed=seed + 7, bit_packed=True)  # warm: code, workspace
    r = s.sample(shots, seed=seed, bit_packed=True)
    s.close()
    return r.kernel_s / shots


def measure_labels(blobs, log=print, every: int = 1000):
    td = np.zeros(len(blobs))
    tc = np.zeros(len(blobs))
    t0 = time.time()
    for i, b in enumerate(blobs):
        td[i] = time_per_shot(b, "dense", seed=1 + i)
        tc[i] = time_per_shot(b, "components", seed=1 + i)
        if (i + 1) % every == 0:
            log(f"labels {i + 1}/{len(blobs)} ({time.time() - t0:.0f}s)")
    return td, tc


def train_model(x_raw, td, tc, epochs: int = 150, batch: int = 64, lr: float = 2e-3, time_weight: float = 0.5,
                seed: int = 7, log=print):
    mean = x_raw.mean(axis=0)
    scale = x_raw.std(axis=0)
    scale[scale < 1e-6] = 1.0
    x = (x_raw - mean) / scale
    y = (tc < td).astype(np.int32)
    lt = np.log10(np.minimum(td, tc))
    t_mean, t_scale = float(lt.mean()), float(lt.std() or 1.0)
    t = (lt - t_mean) / t_scale
    net = DeviceMLP(init_params(seed))
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    steps = n // batch
    history = []
    for ep in range(epochs):
        order = rng.permutation(n)[: steps * batch].astype(np.uint32)
        losses = net.train(x, y, t, order, batch, lr, time_weight)
        history.append(float(losses.mean()))
        if ep % 25 == 0 or ep == epochs - 1:
            log(f"epoch {ep}: mean loss {history[-1]:.4f}")
    return net, dict(mean=mean.astype(np.float32), scale=scale.astype(np.float32), t_mean=t_mean,
                     t_scale=t_scale, history=np.array(history, dtype=np.float32))


def evaluate(p_components, heuristic, td, tc, margin: float) -> dict:
    """Regret of each policy on a batch: total time of its choices over the
    total time of the per-program best layout, minus one."""
    best = np.minimum(td, tc).sum()
    learned, confident = decide(p_components, heuristic, margin)
    cost = lambda choice: float(np.where(choice, tc, td).sum())
    out = dict(best=float(best), learned=cost(learned), heuristic=cost(heuristic.astype(bool)),
               dense=float(td.sum()), components=float(tc.sum()))
    for k in ("learned", "heuristic", "dense", "components"):
        out[f"regret_{k}"] = out[k] / out["best"] - 1.0
    truth = tc < td
    out["accuracy_learned"] = float(np.mean(learned == truth))
    out["accuracy_heuristic"] = float(np.mean(heuristic.astype(bool) == truth))
    out["fallback_fraction"] = float(np.mean(~confident))
    out["gain_over_heuristic"] = out["heuristic"] / out["learned"]
    return out


def main(argv) -> int:
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--heldout", type=int, default=512)
    ap.add_argument("--heldout-start", type=int, default=100000)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--margin", type=float, default=0.1)
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--dataset", default="")
    ap.add_argument("--workers", type=int, default=16)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    t0 = time.time()
    tr_idx, tr_blobs, tr_x = planned_corpus(0, a.train, a.workers)
    ho_idx, ho_blobs, ho_x = planned_corpus(a.heldout_start, a.heldout, a.workers)
    log(f"corpus: {len(tr_idx)} training programs (indices < {tr_idx.max() + 1}), {len(ho_idx)} held out "
        f"({time.time() - t0:.0f}s)")
    heur_tr = np.array([int(b[9]) & 1 for b in tr_blobs], dtype=np.int32)
    heur_ho = np.array([int(b[9]) & 1 for b in ho_blobs], dtype=np.int32)
    td_tr, tc_tr = measure_labels(tr_blobs, log)
    td_ho, tc_ho = measure_labels(ho_blobs, log)
    log(f"labels done ({time.time() - t0:.0f}s); components faster on {np.mean(tc_tr < td_tr):.3f} of training")
    if a.dataset:
        np.savez(a.dataset, tr_idx=tr_idx, tr_x=tr_x, td_tr=td_tr, tc_tr=tc_tr, heur_tr=heur_tr,
                 ho_idx=ho_idx, ho_x=ho_x, td_ho=td_ho, tc_ho=tc_ho, heur_ho=heur_ho)
    net, norm = train_model(tr_x, td_tr, tc_tr, epochs=a.epochs, log=log)
    out = net.predict((ho_x - norm["mean"]) / norm["scale"])
    p = softmax_p_components(out)
    ev = evaluate(p, heur_ho, td_ho, tc_ho, a.margin)
    out_tr = net.predict((tr_x - norm["mean"]) / norm["scale"])
    ev_tr = evaluate(softmax_p_components(out_tr), heur_tr, td_tr, tc_tr, a.margin)
    log("held-out: " + ", ".join(f"{k} {v:.4g}" for k, v in ev.items()))
    log("training: " + ", ".join(f"{k} {v:.4g}" for k, v in ev_tr.items()))
    np.savez(a.out, params=net.params(), mean=norm["mean"], scale=norm["scale"], margin=np.float32(a.margin),
             t_mean=np.float64(norm["t_mean"]), t_scale=np.float64(norm["t_scale"]), history=norm["history"],
             ho_idx=ho_idx, ho_x=ho_x, ho_td=td_ho, ho_tc=tc_ho, ho_heuristic=heur_ho,
             ho_p_components=p.astype(np.float32), n_train=np.int64(len(tr_idx)),
             train_max_index=np.int64(tr_idx.max()))
    log(f"wrote {a.out} ({time.time() - t0:.0f}s)")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
