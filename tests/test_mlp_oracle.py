"""The MLP oracle's gradient against central finite differences (CPU)."""

import numpy as np

from oracle import mlp


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = rng.normal(0, 0.3, mlp.NUM_PARAMS)
    X = rng.normal(size=(40, 8))
    T = rng.normal(size=(40, 3))
    loss, g = mlp.loss_and_grad(theta, X, T)
    for j in rng.choice(mlp.NUM_PARAMS, 60, replace=False):
        h = 1e-6
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fd = (mlp.loss_and_grad(tp, X, T)[0] - mlp.loss_and_grad(tm, X, T)[0]) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 + 1e-5 * abs(fd)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0], np.float32)
    r = mlp.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and abs(r[3] + 3.140625) < 1e-12 and r[4] == 0.0
