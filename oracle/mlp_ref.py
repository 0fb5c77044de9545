"""FP64 CPU restatement of the closure MLP (45 -> 64 -> 64 -> 2, ReLU) with its
forward pass, backward pass (MSE loss) and Adam step.  TEST INFRASTRUCTURE ONLY:
it is the reference for the GPU kernels in sailfish_amd/csrc/mlp.hip and is pinned by
central finite differences in tests/test_mlp_ref.py.

Parameter vector layout (shared with the GPU kernels, 8336 FP32 values):
  W1[48][64] (input rows 45..47 are padding and stay 0), b1[64],
  W2[64][64], b2[64], W3[64][16] (output columns 2..15 are padding), b3[16].
Loss: mean over the batch and the 2 outputs of (y_hat - y)^2.
"""
from __future__ import annotations

import math

import numpy as np

NIN, IN, H, NOUT, OUT = 45, 48, 64, 2, 16
SHAPES = (("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, OUT)), ("b3", (OUT,)))
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(p) -> np.ndarray:
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def init_params(seed: int = 7) -> np.ndarray:
    """He-style initialisation; padding rows/columns zero."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = {
        "W1": rng.standard_normal((IN, H)) * math.sqrt(2.0 / NIN),
        "b1": np.zeros(H),
        "W2": rng.standard_normal((H, H)) * math.sqrt(2.0 / H),
        "b2": np.zeros(H),
        "W3": rng.standard_normal((H, OUT)) * math.sqrt(1.0 / H),
        "b3": np.zeros(OUT),
    }
    p["W1"][NIN:, :] = 0.0
    p["W3"][:, NOUT:] = 0.0
    return pack(p)


def relu(x):
    return np.maximum(x, 0.0)


def relu_grad(x):
    """Derivative of max(x, 0), taken as 0 at x = 0."""
    return (np.asarray(x) > 0.0).astype(np.float64)


def forward(theta, X):
    """X: (B, 48) (columns 45..47 zero).  Returns outputs (B, 2) and the cache."""
    p = unpack(theta)
    Z1 = X @ p["W1"] + p["b1"]
    A1 = relu(Z1)
    Z2 = A1 @ p["W2"] + p["b2"]
    A2 = relu(Z2)
    Z3 = A2 @ p["W3"] + p["b3"]
    return Z3[:, :NOUT], (X, Z1, A1, Z2, A2, Z3)


def loss(theta, X, Y):
    out, _ = forward(theta, X)
    return float(np.mean((out - Y) ** 2))


def grad(theta, X, Y):
    """Gradient of the loss with respect to the parameter vector, and the loss."""
    p = unpack(theta)
    out, (X, Z1, A1, Z2, A2, Z3) = forward(theta, X)
    B = X.shape[0]
    dZ3 = np.zeros_like(Z3)
    dZ3[:, :NOUT] = 2.0 * (out - Y) / (B * NOUT)
    g = {"W3": A2.T @ dZ3, "b3": dZ3.sum(0)}
    dZ2 = (dZ3 @ p["W3"].T) * relu_grad(Z2)
    g["W2"] = A1.T @ dZ2
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ p["W2"].T) * relu_grad(Z1)
    g["W1"] = X.T @ dZ1
    g["b1"] = dZ1.sum(0)
    return pack(g), float(np.mean((out - Y) ** 2))


def adam(theta, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new theta, m, v."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_round(x) -> np.ndarray:
    """Round FP32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# ---- the GPU kernels' rounding, emulated in FP64 (reference for the MFMA kernels) ----
def forward_bf16(theta, X):
    """Forward pass with the GPU's operand rounding: BF16 inputs and weights, exact
    products and sums, FP32-level biases, hidden activations relu(bf16(Z))."""
    p = unpack(theta)
    W1, W2, W3 = bf16_round(p["W1"]), bf16_round(p["W2"]), bf16_round(p["W3"])
    Xb = bf16_round(X)
    Z1 = Xb @ W1 + p["b1"]
    A1 = relu(bf16_round(Z1))
    Z2 = A1 @ W2 + p["b2"]
    A2 = relu(bf16_round(Z2))
    Z3 = A2 @ W3 + p["b3"]
    return Z3[:, :NOUT], (Xb, Z1, A1, Z2, A2, Z3, W1, W2, W3)


def grad_bf16(theta, X, Y):
    """Gradient with the GPU's operand rounding (BF16 dZ operands for the products,
    unrounded dZ for the bias sums; the ReLU derivative is taken on the rounded
    activation, [A > 0])."""
    out, (Xb, Z1, A1, Z2, A2, Z3, W1, W2, W3) = forward_bf16(theta, X)
    B = X.shape[0]
    dZ3 = np.zeros_like(Z3)
    dZ3[:, :NOUT] = 2.0 * (out - Y) / (B * NOUT)
    dZ3b = bf16_round(dZ3)
    g = {"W3": A2.T @ dZ3b, "b3": dZ3.sum(0)}
    dZ2 = (dZ3b @ W3.T) * relu_grad(A2)
    dZ2b = bf16_round(dZ2)
    g["W2"] = A1.T @ dZ2b
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2b @ W2.T) * relu_grad(A1)
    dZ1b = bf16_round(dZ1)
    g["W1"] = Xb.T @ dZ1b
    g["b1"] = dZ1.sum(0)
    return pack(g), float(np.mean((out - Y) ** 2))
