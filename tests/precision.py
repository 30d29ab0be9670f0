"""Error bounds for comparing device probabilities with the oracle (test-only).

The FP64 paths compute the same amplitudes as the oracle operation by
operation; they differ in how a measurement's probability is summed (the
oracle in the reference's scalar index order, the device with a fixed tree)
and, through 1/sqrt(P(branch)), in the scale of every later state. For a
sum of n non-negative terms, recursive summation is within gamma_{n-1} * P
of the exact sum and a binary tree within gamma_{ceil(log2 n)} * P, where
gamma_m = m u / (1 - m u) and u = 2^-53 (Higham, Accuracy and Stability of
Numerical Algorithms, 2nd ed., section 4.2). Collapsing onto a branch of
probability P_b divides by the computed P_b, which leaves the new state's
norm off by at most E / P_b, E the absolute error of that sum; those norm
errors multiply every later probability. So for measurement m of a shot:

    |P_device,m - P_oracle,m| <= 2 * (g * P_m + P_m * sum_{i<m} g * P_i / P_b,i)

with g = gamma_{n-1} + gamma_{ceil(log2 n) + 2} for n = 2^(w-1) terms, w the
widest component measured, and the factor 2 covering the elementwise
rounding of the two scaled states.
"""
from __future__ import annotations

import math

import numpy as np

U = 2.0 ** -53


def gamma(m: float) -> float:
    return m * U / (1.0 - m * U)


def fp64_sum_bound(probs: np.ndarray, uniforms: np.ndarray, width: int) -> np.ndarray:
    """Per (measurement, shot) bound of the formula above; `probs` and
    `uniforms` are [measurements][shots] (the oracle's P(true) and the
    injected draws, branch = u < P)."""
    n = float(2 ** max(0, width - 1))
    g = gamma(max(n - 1.0, 1.0)) + gamma(math.ceil(math.log2(max(n, 2.0))) + 2)
    taken = np.where(uniforms < probs, probs, 1.0 - probs)
    drift = g * probs / np.maximum(taken, 1e-300)
    prior = np.vstack([np.zeros((1, probs.shape[1])), np.cumsum(drift, axis=0)[:-1]])
    return 2.0 * (g * probs + probs * prior)


# The FP32 amplitude path's bar against the FP64 oracle: probabilities (the
# 1e-6 target) and amplitudes.
FP32_PROB_TOL = 1e-6
FP32_AMP_TOL = 1e-6
