"""Seeded initial conditions for the benchmark workloads (oracle side).

The product package has its own generators (``jaccpot_amd.ics``); this copy
belongs to the oracle.  Both follow the same published formulas:

* uniform: U[-1,1]^3, masses U[0.5,1.5];
* Plummer a=1 truncated at r=10a: u ~ U(0, M(10)), M(r) = r^3/(r^2+1)^{3/2},
  r = (u^{-2/3} - 1)^{-1/2}, isotropic direction, masses 1/N;
* Gaussian blobs: 64 centres U[-1,1]^3, members N(c, 0.05^2 I), masses lognormal(0, 0.5).
"""

from __future__ import annotations

import numpy as np


def uniform(n: int, seed: int):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, size=(n, 3))
    m = rng.uniform(0.5, 1.5, size=n)
    return x, m


def plummer(n: int, seed: int, a: float = 1.0, rmax: float = 10.0):
    rng = np.random.default_rng(seed)
    mmax = rmax ** 3 / (rmax ** 2 + 1.0) ** 1.5
    u = rng.uniform(0.0, mmax, size=n)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    cth = rng.uniform(-1.0, 1.0, size=n)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=n)
    sth = np.sqrt(1.0 - cth * cth)
    x = np.stack([r * sth * np.cos(phi), r * sth * np.sin(phi), r * cth], axis=1)
    m = np.full(n, 1.0 / n)
    return x, m


def blobs(n: int, seed: int, nblobs: int = 64, sigma: float = 0.05):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-1.0, 1.0, size=(nblobs, 3))
    which = np.arange(n) % nblobs
    x = centres[which] + sigma * rng.standard_normal(size=(n, 3))
    m = rng.lognormal(0.0, 0.5, size=n)
    return x, m


def galaxy(n: int, seed: int, a: float = 1.0, rmax: float = 100.0, r_d: float = 0.3, z_d: float = 0.03,
           halo_fraction: float = 0.8):
    """Hernquist halo (a, truncated at rmax, r = a sqrt(u)/(1 - sqrt(u)), u ~ U(0, M(rmax)),
    M(r) = r^2/(r+a)^2) holding halo_fraction of the equal-mass particles, plus an
    exponential disc: R ~ Gamma(2, r_d), z ~ Laplace(0, z_d), uniform azimuth."""
    rng = np.random.default_rng(seed)
    nh = int(round(halo_fraction * n))
    nd = n - nh
    umax = rmax * rmax / (rmax + a) ** 2
    u = rng.uniform(0.0, umax, size=nh)
    su = np.sqrt(u)
    r = a * su / (1.0 - su)
    cth = rng.uniform(-1.0, 1.0, size=nh)
    ph = rng.uniform(0.0, 2.0 * np.pi, size=nh)
    sth = np.sqrt(1.0 - cth * cth)
    halo = np.stack([r * sth * np.cos(ph), r * sth * np.sin(ph), r * cth], axis=1)
    R = rng.gamma(2.0, r_d, size=nd)
    phd = rng.uniform(0.0, 2.0 * np.pi, size=nd)
    z = rng.laplace(0.0, z_d, size=nd)
    disc = np.stack([R * np.cos(phd), R * np.sin(phd), z], axis=1)
    x = np.concatenate([halo, disc], axis=0)
    return x, np.full(n, 1.0 / n)
