"""Real (Dehnen no-sqrt2) solid-harmonic operators, float64 NumPy restatement.

Oracle / test infrastructure only (see ``oracle/__init__.py``).

Every function restates the algorithm of the reference file:line it cites, in
the same order of operations where that order is part of the numerics (the
Legendre / Chebyshev recurrences, the per-degree rotation cascade, the
ascending-k z-translation sums).  Inputs are batched over leading axes.

Parity pins (tests/test_oracle_operators.py):
  * U_n^m: Dehnen Table 3 degree 1, the derivative recurrences
    dU_n^m/dz = U_{n-1}^m and the half-integer transverse identities, U_0^0 = 1.
  * B matrices: B^2 = I, checkerboard sparsity.
  * rotations: Dz(0) = I, z-aligned = I, D_to p2m(s) = p2m(g s), to/from inverse.
  * translations: identity at dz = 0, M2L monopole -> 1/r, convergence in p,
    the single-particle convention fixture and the per-direction error table of
    the reference's operator-conventions document.
"""

from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

# ---------------------------------------------------------------------------
# packed layout (jaccpot/operators/_sh_indexing.py:36-112)
# ---------------------------------------------------------------------------


def sh_size(p: int) -> int:
    """(p+1)^2 coefficients up to degree p (_sh_indexing.py:36-57)."""
    if p < 0:
        raise ValueError("order must be >= 0")
    return (p + 1) * (p + 1)


def sh_index(n: int, m: int) -> int:
    """Packed index n^2 + n + m, m = -n..n (_sh_indexing.py:85-112)."""
    if n < 0 or m < -n or m > n:
        raise ValueError("need n >= 0 and -n <= m <= n")
    return n * n + n + m


def factorial_table(max_n: int) -> np.ndarray:
    """n! as exp(lgamma(n+1)) in float64 (_sh_indexing.py:120-146).

    Deliberately the same inexact construction as the reference (not the exact
    integers), so the oracle carries the same ~1e-14 factorial rounding.
    """
    return np.array([math.exp(math.lgamma(n + 1.0)) for n in range(max_n + 1)])


SQUARED_RADIUS_FLOOR_F64 = 1e-60  # operators/dtypes.py:23-55 (normal in float64)


# ---------------------------------------------------------------------------
# forward-mode dual numbers (value + d/dx, d/dy, d/dz) for exact gradients
# ---------------------------------------------------------------------------


class Dual:
    """value ``v`` with its three partial derivatives ``d[0..2]`` (forward AD).

    Used so the oracle's L2P gradient differentiates exactly the same formula
    as its potential, the way the reference takes ``jax.value_and_grad`` of
    ``evaluate_local_real`` (real_p2m_l2p.py:436-440).
    """

    __slots__ = ("v", "d")

    def __init__(self, v, d):
        self.v = np.asarray(v, dtype=np.float64)
        self.d = np.asarray(d, dtype=np.float64)

    @staticmethod
    def const(c, shape):
        v = np.broadcast_to(np.asarray(c, dtype=np.float64), shape).copy()
        return Dual(v, np.zeros((3,) + v.shape))

    def _lift(self, o):
        if isinstance(o, Dual):
            return o
        return Dual(np.broadcast_to(np.asarray(o, np.float64), self.v.shape), np.zeros_like(self.d))

    def __add__(self, o):
        o = self._lift(o)
        return Dual(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __sub__(self, o):
        o = self._lift(o)
        return Dual(self.v - o.v, self.d - o.d)

    def __rsub__(self, o):
        return self._lift(o) - self

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __mul__(self, o):
        o = self._lift(o)
        return Dual(self.v * o.v, self.d * o.v + self.v * o.d)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = self._lift(o)
        q = self.v / o.v
        return Dual(q, (self.d - q * o.d) / o.v)

    def __rtruediv__(self, o):
        return self._lift(o) / self

    def __pow__(self, k: int):
        if k == 0:
            return Dual.const(1.0, self.v.shape)
        out = self
        for _ in range(k - 1):
            out = out * self
        return out

    def sqrt(self):
        s = np.sqrt(self.v)
        return Dual(s, self.d / (2.0 * s))

    def floor_max(self, floor: float):
        """max(self, floor): the derivative is zero wherever the floor binds."""
        bind = self.v < floor
        return Dual(np.where(bind, floor, self.v), np.where(bind, 0.0, self.d))


def _sqrt(a):
    return a.sqrt() if isinstance(a, Dual) else np.sqrt(a)


def _floor(a, floor):
    return a.floor_max(floor) if isinstance(a, Dual) else np.maximum(a, floor)


# ---------------------------------------------------------------------------
# solid harmonics U_n^m (real_p2m_l2p.py:51-232, 241-399)
# ---------------------------------------------------------------------------


def _solid_harmonics_generic(x, y, z, p: int):
    """U_n^m(x, y, z) for all n <= p, packed (list of C entries).

    Follows real_p2m_l2p.py:126-232: floored r and rho, cos/sin theta, the
    Chebyshev recurrences for cos(m phi) / sin(m phi) seeded with x/rho, y/rho
    (_sh_indexing.py:149-195), the Legendre recursion without Condon-Shortley
    phase, and division by (n+|m|)! from the lgamma factorial table.
    Works on ndarrays or on :class:`Dual` values.
    """
    r2 = x * x + y * y + z * z
    r = _sqrt(_floor(r2, SQUARED_RADIUS_FLOOR_F64))
    rho2 = x * x + y * y
    rho = _sqrt(_floor(rho2, SQUARED_RADIUS_FLOOR_F64))
    cos_t = z / r
    sin_t = rho / r
    cos_p = x / rho
    sin_p = y / rho
    fact = factorial_table(2 * p)
    shape = x.v.shape if isinstance(x, Dual) else np.shape(x)

    def const(c):
        return Dual.const(c, shape) if isinstance(x, Dual) else np.full(shape, float(c))

    cos_m = [const(1.0)]
    sin_m = [const(0.0)]
    if p >= 1:
        cos_m.append(cos_p)
        sin_m.append(sin_p)
    for m in range(2, p + 1):
        cos_m.append(2.0 * cos_p * cos_m[m - 1] - cos_m[m - 2])
        sin_m.append(2.0 * cos_p * sin_m[m - 1] - sin_m[m - 2])

    out = [None] * sh_size(p)
    for n in range(p + 1):
        r_n = r ** n if n > 0 else const(1.0)
        for m in range(-n, n + 1):
            ma = abs(m)
            if ma == 0:
                pmm = const(1.0)
            else:
                dfact = fact[2 * ma] / ((2.0 ** ma) * fact[ma])
                pmm = dfact * (sin_t ** ma)
            if ma == n:
                pnm = pmm
            elif ma + 1 == n:
                pnm = (2.0 * ma + 1.0) * cos_t * pmm
            else:
                p2 = pmm
                p1 = (2.0 * ma + 1.0) * cos_t * pmm
                for k in range(ma + 2, n + 1):
                    pk = ((2.0 * k - 1.0) * cos_t * p1 - (k + ma - 1.0) * p2) / (k - ma)
                    p2, p1 = p1, pk
                pnm = p1
            trig = cos_m[ma] if m >= 0 else sin_m[ma]
            out[sh_index(n, m)] = r_n * pnm * trig / fact[n + ma]
    return out


def solid_harmonics(delta: np.ndarray, p: int) -> np.ndarray:
    """U_n^m(delta) packed, shape [..., (p+1)^2] (real_p2m_l2p.py:51-232)."""
    d = np.asarray(delta, dtype=np.float64)
    vals = _solid_harmonics_generic(d[..., 0], d[..., 1], d[..., 2], p)
    return np.stack(vals, axis=-1)


def solid_harmonics_with_grad(delta: np.ndarray, p: int):
    """(U [..., C], dU/d delta [..., C, 3]) by forward-mode AD of the same formula."""
    d = np.asarray(delta, dtype=np.float64)
    shape = d.shape[:-1]
    comps = []
    for a in range(3):
        dv = np.zeros((3,) + shape)
        dv[a] = 1.0
        comps.append(Dual(d[..., a], dv))
    vals = _solid_harmonics_generic(comps[0], comps[1], comps[2], p)
    u = np.stack([v.v for v in vals], axis=-1)
    g = np.stack([np.moveaxis(v.d, 0, -1) for v in vals], axis=-2)
    return u, g


def p2m(deltas: np.ndarray, masses: np.ndarray, p: int) -> np.ndarray:
    """Multipole sum_i m_i U(x_i - c) (real_p2m_l2p.py:51-232, real_tree_expansions.py:175-207).

    ``deltas`` are particle - centre, shape [K, 3]; returns [(p+1)^2].
    """
    u = solid_harmonics(deltas, p)
    return np.einsum("k,kc->c", np.asarray(masses, np.float64), u)


def l2p(local: np.ndarray, deltas: np.ndarray, p: int):
    """phi = sum F_n^m U_n^m(delta) and grad_delta phi, delta = centre - eval point.

    real_p2m_l2p.py:241-399 (potential) and :494-547 (gradient by autodiff of
    the potential).  ``local`` is [C] or [..., C] broadcast against deltas[..., 3].
    Returns (phi [...], grad [..., 3]).
    """
    u, g = solid_harmonics_with_grad(deltas, p)
    F = np.asarray(local, np.float64)
    phi = np.einsum("...c,...c->...", np.broadcast_to(F, u.shape), u)
    grad = np.einsum("...c,...ca->...a", np.broadcast_to(F, u.shape), g)
    return phi, grad


# ---------------------------------------------------------------------------
# Dehnen B and Q matrices (real_dehnen_q.py:49-204, 343-412)
# ---------------------------------------------------------------------------


@lru_cache(maxsize=None)
def dehnen_B_complex(ell: int) -> np.ndarray:
    """Complex swap matrix B_ell by the Dehnen (2014) eq. 63 recursion.

    real_dehnen_q.py:49-145: start B_0 = [[1]]; interior rows by 63a, the top
    edge row m = n+1 by 63b, the bottom edge row m = -(n+1) by 63c.
    """
    B = np.array([[1.0]])

    def get(Bp, n, m, l):
        if abs(m) > n or abs(l) > n:
            return 0.0
        return Bp[m + n, l + n]

    for n in range(ell):
        nn = n + 1
        Bn = np.zeros((2 * nn + 1, 2 * nn + 1))
        for mt in range(-nn, nn + 1):
            for lt in range(-nn, nn + 1):
                if abs(mt) <= n:
                    val = 0.5 * (get(B, n, mt, lt - 1) - get(B, n, mt, lt + 1))
                elif mt == n + 1:
                    val = 0.5 * (get(B, n, n, lt - 1) + get(B, n, n, lt + 1) + 2.0 * get(B, n, n, lt))
                else:
                    val = 0.5 * (get(B, n, -n, lt - 1) + get(B, n, -n, lt + 1) - 2.0 * get(B, n, -n, lt))
                Bn[mt + nn, lt + nn] = val
        B = Bn
    return B


@lru_cache(maxsize=None)
def dehnen_Q(ell: int) -> np.ndarray:
    """Complex -> Dehnen no-sqrt2 real map for one degree (real_dehnen_q.py:146-200).

    T_{+m} = (Th_m + (-1)^m Th_{-m}) / 2,  T_{-m} = -i (Th_m - (-1)^m Th_{-m}) / 2.
    """
    n = 2 * ell + 1
    Q = np.zeros((n, n), dtype=np.complex128)
    Q[ell, ell] = 1.0
    for m in range(1, ell + 1):
        ph = (-1.0) ** m
        Q[ell + m, ell + m] = 0.5
        Q[ell + m, ell - m] = 0.5 * ph
        Q[ell - m, ell + m] = -0.5j
        Q[ell - m, ell - m] = 0.5j * ph
    return Q


@lru_cache(maxsize=None)
def real_B_matrices(ell: int):
    """(B_T, B_U) = Re(Q B Q^-1), Re(Q B^T Q^-1) (real_dehnen_q.py:343-412)."""
    Bc = dehnen_B_complex(ell)
    Q = dehnen_Q(ell)
    Qi = np.linalg.inv(Q)
    BT = Q @ Bc @ Qi
    BU = Q @ Bc.T @ Qi
    return np.real(BT).copy(), np.real(BU).copy()


def B_U(ell: int) -> np.ndarray:
    """Multipole swap matrix used by every rotation (real_dehnen_q.py:443-470)."""
    return real_B_matrices(ell)[1]


# ---------------------------------------------------------------------------
# rotations (real_rotations.py:75-147, 340-501, 541-649)
# ---------------------------------------------------------------------------


def Dz(ell: int, angle) -> np.ndarray:
    """z-rotation block [..., 2l+1, 2l+1] (real_rotations.py:75-147)."""
    a = np.asarray(angle, np.float64)
    n = 2 * ell + 1
    D = np.zeros(a.shape + (n, n))
    D[..., ell, ell] = 1.0
    for m in range(1, ell + 1):
        c = np.cos(m * a)
        s = np.sin(m * a)
        ip, im = ell + m, ell - m
        D[..., ip, ip] = c
        D[..., ip, im] = -s
        D[..., im, ip] = s
        D[..., im, im] = c
    return D


def alignment_angles(x, y, z):
    """az = atan2(x, y) (0 at rho=0), ax = atan2(rho, z) (0 at r=0) (real_rotations.py:340-414)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    rho_sq = x * x + y * y
    rho_pos = rho_sq > 0
    rho = np.where(rho_pos, np.sqrt(np.where(rho_pos, rho_sq, 1.0)), 0.0)
    az = np.where(rho_pos, np.arctan2(np.where(rho_pos, x, 1.0), y), 0.0)
    r_pos = (rho_sq + z * z) > 0
    ax = np.where(r_pos, np.arctan2(rho, np.where(r_pos, z, 1.0)), 0.0)
    return az, ax


def rot_to_z_multipole(delta: np.ndarray, ell: int) -> np.ndarray:
    """D_to^M = B_U Dz(-ax) B_U Dz(az) (real_rotations.py:417-464)."""
    d = np.asarray(delta, np.float64)
    az, ax = alignment_angles(d[..., 0], d[..., 1], d[..., 2])
    B = B_U(ell)
    return B @ Dz(ell, -ax) @ B @ Dz(ell, az)


def rot_from_z_multipole(delta: np.ndarray, ell: int) -> np.ndarray:
    """D_from^M = Dz(-az) B_U Dz(ax) B_U (real_rotations.py:467-501)."""
    d = np.asarray(delta, np.float64)
    az, ax = alignment_angles(d[..., 0], d[..., 1], d[..., 2])
    B = B_U(ell)
    return Dz(ell, -az) @ B @ Dz(ell, ax) @ B


def rot_from_z_local(delta, ell):
    """Local z->world = (D_to^M)^T (real_rotations.py:541-576)."""
    return np.swapaxes(rot_to_z_multipole(delta, ell), -1, -2)


def rot_to_z_local(delta, ell):
    """Local world->z = (D_from^M)^T (real_rotations.py:614-649)."""
    return np.swapaxes(rot_from_z_multipole(delta, ell), -1, -2)


def _apply_blocks(coeffs: np.ndarray, delta: np.ndarray, p: int, builder) -> np.ndarray:
    c = np.asarray(coeffs, np.float64)
    out = np.zeros_like(c)
    for ell in range(p + 1):
        sl = slice(ell * ell, (ell + 1) * (ell + 1))
        D = builder(delta, ell)
        out[..., sl] = np.einsum("...ij,...j->...i", D, c[..., sl])
    return out


# ---------------------------------------------------------------------------
# z-axis translations (real_translations.py:72-177, 213-367)
# ---------------------------------------------------------------------------


def zshift(coeffs: np.ndarray, dz, p: int, which: str) -> np.ndarray:
    """Same-type z shift, summed over k ascending (real_translations.py:72-177).

    m2m: out_n^m = sum_{k=0}^{n-|m|} dz^k/k! c_{n-k}^m
    l2l: out_n^m = sum_{k=0}^{p-n}   dz^k/k! c_{n+k}^m
    """
    c = np.asarray(coeffs, np.float64)
    dz = np.asarray(dz, np.float64)
    fact = factorial_table(p)
    ck = [dz ** k / fact[k] for k in range(p + 1)]
    out = np.zeros_like(c)
    for n in range(p + 1):
        for m in range(-n, n + 1):
            acc = np.zeros(c.shape[:-1])
            for k in range(p + 1):
                src_n = n - k if which == "m2m" else n + k
                ok = src_n >= abs(m) if which == "m2m" else src_n <= p
                if ok:
                    acc = acc + c[..., sh_index(src_n, m)] * ck[k]
            out[..., sh_index(n, m)] = acc
    return out


def zm2l(coeffs: np.ndarray, r, p: int) -> np.ndarray:
    """z-axis M2L (real_translations.py:213-334, floor 1e-30 at m2l_real_rot_scale.py:322).

    F_n^m = sum_{k=|m|}^{p-n} s(m) (n+k)! / r^{n+k+1} M_k^m,  s(m) = (-1)^m (2 if m != 0 else 1).
    """
    c = np.asarray(coeffs, np.float64)
    r = np.maximum(np.asarray(r, np.float64), 1.0e-30)
    fact = factorial_table(2 * p)
    out = np.zeros_like(c)
    for n in range(p + 1):
        for m in range(-n, n + 1):
            s = (-1.0 if (m % 2) else 1.0) * (2.0 if m != 0 else 1.0)
            acc = np.zeros(c.shape[:-1])
            for k in range(abs(m), p - n + 1):
                acc = acc + s * fact[n + k] / r ** (n + k + 1) * c[..., sh_index(k, m)]
            out[..., sh_index(n, m)] = acc
    return out


# ---------------------------------------------------------------------------
# full translation operators (real_translations.py:381-667)
# ---------------------------------------------------------------------------


def _radius(delta):
    d = np.asarray(delta, np.float64)
    r2 = np.sum(d * d, axis=-1)
    return np.sqrt(np.maximum(r2, SQUARED_RADIUS_FLOOR_F64))


def m2m(multipole: np.ndarray, delta: np.ndarray, p: int) -> np.ndarray:
    """Multipole re-centring, delta = source(child) - destination(parent).

    real_translations.py:528-598: rotate to z (D_to^M), z-shift by |delta|
    (0 when |delta| < 1e-30), rotate back with D_from^M.
    """
    r = _radius(delta)
    dz = np.where(r < 1e-30, 0.0, r)
    mz = _apply_blocks(multipole, delta, p, rot_to_z_multipole)
    mz = zshift(mz, dz, p, "m2m")
    return _apply_blocks(mz, delta, p, rot_from_z_multipole)


def l2l(local: np.ndarray, delta: np.ndarray, p: int) -> np.ndarray:
    """Local re-centring, delta = parent - child (real_translations.py:601-667)."""
    r = _radius(delta)
    dz = np.where(r < 1e-30, 0.0, r)
    lz = _apply_blocks(local, delta, p, rot_to_z_local)
    lz = zshift(lz, dz, p, "l2l")
    return _apply_blocks(lz, delta, p, rot_from_z_local)


def m2l(multipole: np.ndarray, delta: np.ndarray, p: int) -> np.ndarray:
    """Multipole -> local, delta = target - source (real_translations.py:381-449,
    batched form m2l_real_rot_scale.py:330-375)."""
    d = np.asarray(delta, np.float64)
    r2 = np.sum(d * d, axis=-1)
    r = np.where(r2 > 0, np.sqrt(np.where(r2 > 0, r2, 1.0)), 0.0)
    mz = _apply_blocks(multipole, d, p, rot_to_z_multipole)
    lz = zm2l(mz, r, p)
    return _apply_blocks(lz, d, p, rot_from_z_local)


def truncate(coeffs: np.ndarray, p_from: int, p_to: int) -> np.ndarray:
    """Keep degrees <= p_to of a packed order-p_from vector."""
    assert p_to <= p_from
    return np.asarray(coeffs)[..., : sh_size(p_to)]


def m2l_flops(p: int) -> int:
    """Algorithmic flops per M2L pair: 4 S(p) + 2 Z(p).

    S(p) = sum_{l<=p} (2l+1)^2 counts the checkerboard-sparse rotation
    applications (two per side, 2 flops per non-zero, half the entries zero);
    Z(p) counts the non-zero z-core terms.  F(2)=156, F(4)=718, F(6)=1964.
    """
    S = (p + 1) * (2 * p + 1) * (2 * p + 3) // 3

    def T(k):
        return k * (k + 1) // 2

    Z = T(p + 1) + 2 * sum(T(p - 2 * m + 1) for m in range(1, p // 2 + 1))
    return 4 * S + 2 * Z
