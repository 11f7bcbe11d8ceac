"""CPU: the pair-label oracle against closed forms and invariances."""

from __future__ import annotations

import numpy as np

from oracle import harmonics as H
from oracle import labels as OL


def _rotation(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_point_mass_closed_form_any_direction():
    """A unit point mass at the source centre: N_n(L) = 1 / d^(n+1), so
    e_g = sqrt(sum_{n>g} (n r^(n-1) / d^(n+1))^2) / sqrt(sum_{n>=1} (...)^2)."""
    rng = np.random.default_rng(0)
    d, r_t = 2.7, 0.9
    n = np.arange(9)
    a = np.where(n > 0, n * r_t ** np.maximum(n - 1, 0) / d ** (n + 1), 0.0)
    expected = [np.log10(np.sqrt(np.sum(a[g + 1 :] ** 2)) / np.sqrt(np.sum(a ** 2))) for g in (2, 4, 6)]
    M8 = np.zeros((1, H.sh_size(8)))
    M8[0, 0] = 1.0
    for _ in range(5):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        got = OL.pair_labels(M8, (d * u)[None], np.array([r_t]))
        np.testing.assert_allclose(got[0], expected, atol=1e-11)
    # a plain (unweighted) coefficient norm, or potential weights r^n, would differ
    b = r_t ** n / d ** (n + 1)
    wrong = np.log10(np.sqrt(np.sum(b[3:] ** 2)) / np.sqrt(np.sum(b ** 2)))
    assert abs(wrong - expected[0]) > 0.05


def test_degree_norm_is_rotation_invariant():
    rng = np.random.default_rng(1)
    src = rng.normal(size=(40, 3)) * 0.3
    m = rng.uniform(0.5, 1.5, 40)
    M = H.p2m(src, m, 8)[None]
    d = np.array([[1.9, -0.7, 2.2]])
    w = OL.degree_weights(np.array([0.6]), 8)
    base = OL.weighted_norm(H.m2l(M, d, 8), w, 8)
    lab = OL.pair_labels(M, d, np.array([0.6]))
    for _ in range(3):
        R = _rotation(rng)
        Mr = H.p2m(src @ R.T, m, 8)[None]
        dr = d @ R.T
        np.testing.assert_allclose(OL.weighted_norm(H.m2l(Mr, dr, 8), w, 8), base, rtol=1e-10)
        np.testing.assert_allclose(OL.pair_labels(Mr, dr, np.array([0.6])), lab, atol=1e-9)


def test_top_order_has_no_error_and_errors_fall_with_gear():
    rng = np.random.default_rng(2)
    src = rng.normal(size=(30, 3)) * 0.25
    M = H.p2m(src, rng.uniform(0.5, 1.5, 30), 8)[None]
    d = np.array([[0.4, 1.1, -2.0]])
    assert OL.pair_labels(M, d, np.array([0.5]), gears=(8,))[0, 0] == np.log10(OL.LABEL_FLOOR)
    lab = OL.pair_labels(M, d, np.array([0.5]), gears=(2, 4, 6))[0]
    assert lab[0] > lab[1] > lab[2]
