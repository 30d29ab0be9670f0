"""The shipped planner model (ticit_amd/planner_model.npz), checked on the
CPU: the decisions it makes on the held-out batch are reproduced by the
oracle's forward pass, and its measured regret is lower than the reference
heuristic's (the gain of the learned planner over the method without it).

The held-out batch is 512 planned programs from corpus indices >= 100000,
disjoint from the training programs; its per-shot times for both layouts
were measured on an MI355X when the model was trained and are stored with
the model.
"""
import numpy as np

from oracle import mlp
from ticit_amd import planner as PL


def test_model_shape_and_held_out_split():
    m = PL.load_model()
    assert m["params"].shape == (mlp.N_PARAMS,) and m["mean"].shape == (12,) and m["scale"].shape == (12,)
    assert int(m["n_train"]) == 20000 and m["ho_idx"].size == 512
    assert int(m["train_max_index"]) < int(m["ho_idx"].min())


def test_oracle_forward_reproduces_the_device_decisions():
    m = PL.load_model()
    x = (m["ho_x"] - m["mean"]) / m["scale"]
    out = mlp.predict(m["params"], x)
    p = PL.softmax_p_components(out)
    assert np.max(np.abs(p - m["ho_p_components"])) <= 1e-4


def test_learned_planner_beats_the_heuristic_on_held_out_programs():
    m = PL.load_model()
    x = (m["ho_x"] - m["mean"]) / m["scale"]
    p = PL.softmax_p_components(mlp.predict(m["params"], x))
    ev = PL.evaluate(p, m["ho_heuristic"], m["ho_td"], m["ho_tc"], float(m["margin"]))
    print({k: round(v, 5) for k, v in ev.items()})
    assert ev["regret_learned"] < ev["regret_heuristic"]
    assert ev["regret_learned"] < 0.02
    assert ev["gain_over_heuristic"] > 1.0
    assert ev["accuracy_learned"] > ev["accuracy_heuristic"]


def test_the_features_stored_are_the_features_computed():
    from ticit_amd import plan_text
    from ticit_amd.workloads import planner_corpus_circuit
    m = PL.load_model()
    for j in (0, 7, 200):
        f = PL.feature_vector(plan_text(planner_corpus_circuit(int(m["ho_idx"][j]))))
        assert np.array_equal(f, m["ho_x"][j])
