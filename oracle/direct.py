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
