"""Pins of the FP64 CPU MLP (oracle/mlp_ref.py): central finite differences of the
loss match the analytic gradient; the activation (sigmoid form of GELU) matches its
closed form at known points and stays close to x Phi(x);
the bf16 rounding helper is exact on representable values and rounds to nearest even."""
from __future__ import annotations

import math

import numpy as np

from oracle import mlp_ref as R


def _data(B=16, seed=3):
    rng = np.random.default_rng(seed)
    X = np.zeros((B, R.IN))
    X[:, :R.NIN] = rng.standard_normal((B, R.NIN))
    Y = rng.standard_normal((B, R.NOUT))
    return X, Y


def test_gradient_matches_central_differences():
    X, Y = _data()
    th = R.init_params(11)
    th += 0.05 * np.random.default_rng(5).standard_normal(th.shape)   # non-zero biases too
    p = R.unpack(th)
    p["W1"][R.NIN:] = 0.0
    p["W3"][:, R.NOUT:] = 0.0
    th = R.pack(p)
    g, _ = R.grad(th, X, Y)
    rng = np.random.default_rng(9)
    idx = rng.choice(th.size, 200, replace=False)
    h = 1e-6
    for k in idx:
        e = np.zeros_like(th)
        e[k] = h
        fd = (R.loss(th + e, X, Y) - R.loss(th - e, X, Y)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])


def test_padding_gradients_are_zero():
    X, Y = _data()
    g, _ = R.grad(R.init_params(1), X, Y)
    G = R.unpack(g)
    assert np.all(G["W1"][R.NIN:] == 0.0)
    assert np.all(G["W3"][:, R.NOUT:] == 0.0) and np.all(G["b3"][R.NOUT:] == 0.0)


def test_gelu_closed_forms():
    assert R.gelu(np.array([0.0]))[0] == 0.0
    x = 1.0
    assert abs(R.gelu(np.array([x]))[0] - 1.0 / (1.0 + math.exp(-1.702))) < 1e-15
    # the sigmoid form stays within 0.021 of x * Phi(x) (Phi the standard normal CDF)
    xs = np.linspace(-6, 6, 1201)
    phi = 0.5 * (1.0 + np.vectorize(math.erf)(xs / math.sqrt(2.0)))
    assert np.abs(R.gelu(xs) - xs * phi).max() < 0.021
    # gelu(x) - gelu(-x) = x
    xs = np.linspace(-4, 4, 17)
    assert np.allclose(R.gelu(xs) - R.gelu(-xs), xs, atol=1e-14)
    # derivative by central differences
    h = 1e-6
    assert np.allclose((R.gelu(xs + h) - R.gelu(xs - h)) / (2 * h), R.gelu_grad(xs), atol=1e-8)


def test_bf16_rounding():
    vals = np.array([1.0, -2.5, 0.15625, 2.0 ** 100], dtype=np.float32)   # exactly representable in bf16
    assert np.array_equal(R.bf16_round(vals), vals.astype(np.float64))
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert R.bf16_round(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert R.bf16_round(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def test_adam_first_step_is_sign_step():
    th = np.array([1.0, -1.0, 0.5])
    g = np.array([0.3, -2.0, 0.0])
    t1, m, v = R.adam(th, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # bias-corrected first step: delta = lr * g/|g| (eps negligible), 0 for g = 0
    assert np.allclose(t1 - th, -0.01 * np.sign(g), atol=1e-9)
