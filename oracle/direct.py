"""Float64 O(N^2) direct-sum oracle (oracle / test infrastructure only).

Restates jaccpot/autodiff.py:12-98 (``direct_sum_gravitational_acceleration``):
softened Newtonian acceleration with the self term masked,
    a_i = -G sum_{j != i} m_j (x_i - x_j) (|x_i - x_j|^2 + eps^2)^(-3/2),
    phi_i = -G sum_{j != i} m_j (|x_i - x_j|^2 + eps^2)^(-1/2).
Pinned by closed forms in tests (two bodies, Newton's third law, symmetric
shells).
"""

from __future__ import annotations

import numpy as np


def direct_sum(positions, masses, *, G: float = 1.0, eps: float = 0.0, targets=None, chunk: int = 256,
               return_potential: bool = False):
    """Accelerations (and potentials) at ``targets`` (default: all) from every particle."""
    x = np.asarray(positions, np.float64)
    m = np.asarray(masses, np.float64)
    tidx = np.arange(x.shape[0]) if targets is None else np.asarray(targets, np.int64)
    eps2 = float(eps) * float(eps)
    acc = np.zeros((tidx.shape[0], 3))
    pot = np.zeros(tidx.shape[0])
    for s in range(0, tidx.shape[0], chunk):
        t = tidx[s : s + chunk]
        d = x[t][:, None, :] - x[None, :, :]
        r2 = (d * d).sum(-1) + eps2
        mask = t[:, None] != np.arange(x.shape[0])[None, :]
        inv = np.where(mask, 1.0 / np.sqrt(np.where(mask, r2, 1.0)), 0.0)
        w = m[None, :] * inv
        pot[s : s + chunk] = -G * w.sum(axis=1)
        acc[s : s + chunk] = -G * ((w * inv * inv)[..., None] * d).sum(axis=1)
    if return_potential:
        return acc, pot
    return acc


def relative_errors(approx: np.ndarray, exact: np.ndarray) -> np.ndarray:
    """Per-target relative force error |a - a_ref| / |a_ref|."""
    num = np.linalg.norm(np.asarray(approx, np.float64) - exact, axis=-1)
    den = np.linalg.norm(exact, axis=-1)
    return num / np.where(den > 0, den, 1.0)


def direct_sum_vjp(positions, masses, acc_bar, pot_bar, *, G: float = 1.0, eps: float = 0.0, chunk: int = 256):
    """(positions_bar, masses_bar) of sum_i (acc_bar_i . a_i + pot_bar_i phi_i) for the direct sum.

    Closed form of the reverse rule of the softened pair sum (the tidal tensor
    J = -G m (I/r^3 - 3 r r^T / r^5) of nearfield/grad.py:381-389), with
    d = x_j - x_i, s = (|d|^2 + eps^2)^-1/2:
        xbar_i = G sum_j [(I s^3 - 3 d d^T s^5)(m_i abar_j - m_j abar_i) - d s^3 (m_j phib_i + m_i phib_j)]
        mbar_i = G sum_j [-d s^3 . abar_j - s phib_j]
    Pinned by central finite differences (tests/test_oracle_fmm.py).
    """
    x = np.asarray(positions, np.float64)
    m = np.asarray(masses, np.float64)
    ab = np.asarray(acc_bar, np.float64)
    pb = np.asarray(pot_bar, np.float64)
    n = x.shape[0]
    eps2 = float(eps) * float(eps)
    gx = np.zeros((n, 3))
    gm = np.zeros(n)
    for s0 in range(0, n, chunk):
        i = np.arange(s0, min(n, s0 + chunk))
        d = x[None, :, :] - x[i][:, None, :]
        r2 = (d * d).sum(-1) + eps2
        mask = i[:, None] != np.arange(n)[None, :]
        sv = np.where(mask, 1.0 / np.sqrt(np.where(mask, r2, 1.0)), 0.0)
        s3 = sv ** 3
        s5 = s3 * sv * sv
        u = m[i][:, None, None] * ab[None, :, :] - m[None, :, None] * ab[i][:, None, :]
        du = (d * u).sum(-1)
        pw = m[None, :] * pb[i][:, None] + m[i][:, None] * pb[None, :]
        gx[i] = G * (s3[..., None] * u - 3.0 * (s5 * du)[..., None] * d - (s3 * pw)[..., None] * d).sum(1)
        gm[i] = G * (-(s3[..., None] * d * ab[None, :, :]).sum(-1) - sv * pb[None, :]).sum(1)
    return gx, gm
