"""Pins of the FP64 CPU MLP (oracle/mlp_ref.py): central finite differences of the
loss match the analytic gradient; closed forms of the ReLU network (affine when no unit
is cut, bias-only when every layer-1 unit is cut) and of the BF16 rounding model;
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


def _net(seed, scale_b=0.1, nonneg=False):
    rng = np.random.default_rng(seed)
    p = R.unpack(R.init_params(seed))
    for k in ("b1", "b2"):
        p[k] = scale_b * rng.standard_normal(p[k].shape)
    p["b3"][:R.NOUT] = scale_b * rng.standard_normal(R.NOUT)
    if nonneg:
        for k in p:
            p[k] = np.abs(p[k])
    p["W1"][R.NIN:] = 0.0
    p["W3"][:, R.NOUT:] = 0.0
    p["b3"][R.NOUT:] = 0.0
    return p


def test_relu_network_closed_forms():
    # the activation itself
    xs = np.array([-2.0, -1e-300, 0.0, 1e-300, 3.5])
    assert np.array_equal(R.relu(xs), [0.0, 0.0, 0.0, 1e-300, 3.5])
    assert np.array_equal(R.relu_grad(xs), [0.0, 0.0, 0.0, 1.0, 1.0])
    # non-negative inputs, weights and biases: no unit is cut, the network is affine
    p = _net(3, nonneg=True)
    X = np.abs(_data(8)[0])
    out, _ = R.forward(R.pack(p), X)
    lin = ((X @ p["W1"] + p["b1"]) @ p["W2"] + p["b2"]) @ p["W3"] + p["b3"]
    assert np.allclose(out, lin[:, :R.NOUT], rtol=1e-13, atol=0)
    # every layer-1 unit cut (b1 very negative): the output is relu(b2) W3 + b3 for every
    # sample, and nothing flows back into W1, b1
    p = _net(4, scale_b=1.0)
    p["b1"][:] = -1e6
    X, Y = _data(8)
    out, _ = R.forward(R.pack(p), X)
    ref = np.maximum(p["b2"], 0.0) @ p["W3"] + p["b3"]
    assert np.allclose(out, np.broadcast_to(ref[:R.NOUT], out.shape), rtol=1e-14, atol=0)
    assert (p["b2"] < 0).any() and (p["b2"] > 0).any()
    g, _ = R.grad(R.pack(p), X, Y)
    G = R.unpack(g)
    assert np.all(G["W1"] == 0.0) and np.all(G["b1"] == 0.0)
    # the GPU rounding model: activations are relu(bf16(Z)) = bf16(relu(Z)), and the
    # derivative mask is taken on those activations
    z = np.random.default_rng(6).standard_normal(1000) * 3.0
    assert np.array_equal(R.relu(R.bf16_round(z)), R.bf16_round(R.relu(z)))
    pb = _net(5, nonneg=True)
    Xb = R.bf16_round(np.abs(_data(8)[0]))
    outb, cache = R.forward_bf16(R.pack(pb), Xb)
    A1 = cache[2]
    W1, W2, W3 = (R.bf16_round(pb[k]) for k in ("W1", "W2", "W3"))
    assert np.array_equal(A1, R.bf16_round(Xb @ W1 + pb["b1"]))
    A2 = R.bf16_round(A1 @ W2 + pb["b2"])
    assert np.allclose(outb, (A2 @ W3 + pb["b3"])[:, :R.NOUT], rtol=1e-14, atol=0)


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
