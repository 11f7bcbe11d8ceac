"""Float64 oracle of the per-pair truncation-error labels (test infrastructure only).

The labels are this library's own training targets for the learned M2L order
(the reference has no counterpart).  For a far pair (target t, source s) with
displacement d = c_t - c_s, the order-8 translation L8 = M2L_8(M_s) is the
reference and the gear-g translation Lg = M2L_g(M_s truncated to degree g) is
the candidate; the label is

    e_g = sqrt(sum_n w_n^2 N_n(L8 - Lg)^2) / sqrt(sum_n w_n^2 N_n(L8)^2),
    w_n = n r_t^(n-1),
    N_n(L)^2 = sum_m c_nm (L_n^m)^2,  c_nm = 1 / ((n+|m|)! (n-|m|)!) x (1 if m = 0 else 1/2),

the relative size of the pair's force error on the target sphere of radius
r_t: N_n is the rotation-invariant size of the degree-n part of a local
expansion in this real basis (the L2 norm of sum_m L_n^m U_n^m over the unit
sphere, up to a per-degree constant), and w_n the size of a degree-n
gradient at radius r_t.  Stored as log10(max(e_g, 1e-7)).  Pinned by the
closed form of a point mass (N_n = 1 / d^(n+1) in every direction), the
rotation invariance of N_n, and e_8 = 0 (tests/test_learned_oracle.py).
"""

from __future__ import annotations

import numpy as np

from . import harmonics as H

LABEL_FLOOR = 1e-7


def degree_weights(r_t: np.ndarray, p: int = 8) -> np.ndarray:
    """[K, p+1] weights n r_t^(n-1) (zero for n = 0)."""
    r = np.asarray(r_t, np.float64)[:, None]
    n = np.arange(p + 1, dtype=np.float64)[None, :]
    return np.where(n > 0, n * np.power(r, np.maximum(n - 1, 0)), 0.0)


def degree_norm_weights(p: int = 8) -> np.ndarray:
    """[(p+1)^2] packed c_nm = 1 / ((n+|m|)! (n-|m|)!) x (1 if m = 0 else 1/2)."""
    f = H.factorial_table(2 * p)
    c = np.zeros(H.sh_size(p))
    for n in range(p + 1):
        for m in range(-n, n + 1):
            c[H.sh_index(n, m)] = (1.0 if m == 0 else 0.5) / (f[n + abs(m)] * f[n - abs(m)])
    return c


def weighted_norm(L: np.ndarray, w: np.ndarray, p: int) -> np.ndarray:
    """sqrt(sum_n w_n^2 N_n(L)^2) for packed order-p rows L [K, (p+1)^2]."""
    c = degree_norm_weights(p)
    out = np.zeros(L.shape[0])
    for n in range(p + 1):
        sl = slice(n * n, (n + 1) * (n + 1))
        blk = L[:, sl]
        out += w[:, n] ** 2 * np.sum(c[sl] * blk * blk, axis=1)
    return np.sqrt(out)


def pair_labels(M8: np.ndarray, delta: np.ndarray, r_t: np.ndarray, gears=(2, 4, 6)) -> np.ndarray:
    """[K, len(gears)] log10 labels from order-8 source multipoles M8 [K, 81]."""
    M8 = np.asarray(M8, np.float64)
    d = np.asarray(delta, np.float64)
    L8 = H.m2l(M8, d, 8)
    w = degree_weights(r_t, 8)
    ref = np.maximum(weighted_norm(L8, w, 8), 1e-300)
    out = np.zeros((M8.shape[0], len(gears)))
    for j, g in enumerate(gears):
        Lg = np.zeros_like(L8)
        Lg[:, : H.sh_size(g)] = H.m2l(M8[:, : H.sh_size(g)], d, g)
        e = weighted_norm(L8 - Lg, w, 8) / ref
        out[:, j] = np.log10(np.maximum(e, LABEL_FLOOR))
    return out
