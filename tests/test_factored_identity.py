"""The factored measurement identity the wide tier's expectation pass uses
(CPU, mathematics): for a product state |f0> (x) |f1> (x) ... and a measured
body with masks x, z and phase i^ph, the measurement kernels' sum over the
output half (src/contiguous.rs:1149-1320 in the conventions of the device
tables: insert_zero_bit pivots, dir = +-1 by parity(s0 & z), c1 =
conj(i^ph)/sqrt2; diagonal bodies by diagonal_source) equals

    P(true) = (N - Re(i^-ph T)) / 2,
    N = prod_f sum_b |f[b]|^2,
    T = prod_f sum_b (-1)^|b & z_f| conj(f[b]) f[b ^ x_f].

Brute force on random product states of 2-9 qubits; a dropped conjugate, a
wrong sign or a transposed mask field fails it.
"""
import numpy as np
import pytest


def izb(v, bit):
    return ((v >> bit) << (bit + 1)) | (v & ((1 << bit) - 1))


def par(v):
    return bin(v).count("1") & 1


def direct(alpha, W, x, z, piv, ph, diag, dp):
    P = 0.0
    if diag:
        zwop = z & ~(1 << piv)
        for idx in range(2 ** (W - 1)):
            wo = izb(idx, piv)
            pv = dp ^ (par(wo & zwop) if zwop else 0) ^ 1
            P += abs(alpha[wo | (pv << piv)]) ** 2
        return P
    c1 = np.conj(1j ** ph) / np.sqrt(2)
    for idx in range(2 ** (W - 1)):
        s0 = izb(idx, piv)
        d = 1.0 if par(s0 & z) else -1.0
        P += abs(alpha[s0] / np.sqrt(2) + d * c1 * alpha[s0 ^ x]) ** 2
    return P


def factored(fs, ks, x, z, ph):
    N, T, sh = 1.0, 1.0 + 0j, 0
    for k, f in zip(ks, fs):
        m = (1 << k) - 1
        xf, zf = (x >> sh) & m, (z >> sh) & m
        N *= float(np.sum(np.abs(f) ** 2))
        T *= sum((-1) ** par(b & zf) * np.conj(f[b]) * f[b ^ xf] for b in range(1 << k))
        sh += k
    return 0.5 * (N - (1j ** (-ph) * T).real)


@pytest.mark.parametrize("seed", range(6))
def test_factored_probability_equals_direct_sum(seed):
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(40):
        ks = [int(k) for k in rng.integers(1, 4, size=int(rng.integers(2, 4)))]
        W = sum(ks)
        fs = []
        for k in ks:
            f = rng.normal(size=1 << k) + 1j * rng.normal(size=1 << k)
            fs.append(f / np.linalg.norm(f) * rng.uniform(0.5, 1.5))  # norms need not be 1
        alpha = fs[0]
        for f in fs[1:]:
            alpha = np.kron(f, alpha)  # a merge puts the source above the target
        diag = rng.random() < 0.4
        if diag:
            z = int(rng.integers(1, 1 << W))
            x, piv, dp = 0, z.bit_length() - 1, int(rng.integers(2))
            ph = 2 * dp
        else:
            x = int(rng.integers(1, 1 << W))
            z = int(rng.integers(0, 1 << W))
            piv, dp = x.bit_length() - 1, 0
            ph = (2 * int(rng.integers(2)) + par(x & z)) % 4  # Hermitian body
        d = direct(alpha, W, x, z, piv, ph, diag, dp)
        worst = max(worst, abs(d - factored(fs, ks, x, z, ph)))
    assert worst <= 1e-13
