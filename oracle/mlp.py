"""Float64 NumPy oracle of the per-pair order MLP (test infrastructure only).

8 -> 64 -> 64 -> 3 with ReLU; parameters flat as W1[64][8] b1[64] W2[64][64]
b2[64] W3[3][64] b3[3]; the kernels compute with bfloat16-rounded weights, so
the oracle rounds the float32 masters the same way (round to nearest even)
and then works in float64.  Loss = mean over pairs and outputs of the squared
error.  Pinned by central finite differences of ``loss`` (tests).
"""

from __future__ import annotations

import numpy as np

IN, HID, OUT = 8, 64, 3
SIZES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (OUT, HID)), ("b3", (OUT,))]
NUM_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta):
    out = {}
    o = 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o : o + n], np.float64).reshape(shape)
        o += n
    return out


def bf16_round(x):
    """float32 -> bfloat16 (round to nearest even) -> float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward(theta, X):
    p = unpack(theta)
    X = np.asarray(X, np.float64)
    a1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ p["W3"].T + p["b3"]
    return y, (X, a1, h1, a2, h2)


def loss_and_grad(theta, X, T):
    p = unpack(theta)
    y, (X, a1, h1, a2, h2) = forward(theta, X)
    K = X.shape[0]
    e = y - np.asarray(T, np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 * e / (K * OUT)
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    d2 = (dy @ p["W3"]) * (a2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ p["W2"]) * (a1 > 0)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    flat = np.concatenate([g[name].ravel() for name, _ in SIZES])
    return loss, flat
