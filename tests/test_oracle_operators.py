"""Pins of the oracle's operator algebra (CPU).

Each check is chosen so that a plausible mistake fails it: a dropped term,
a sign or index slip, a transposed operand.  Known answers come from the
reference's own tests and documents (cited), closed forms, and exactness
identities of the expansions (M2M and L2L are exact for the degrees kept).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import harmonics as H

RNG = np.random.default_rng(20261202)


def test_packed_layout():
    # _sh_indexing.py:36-112
    assert H.sh_size(0) == 1 and H.sh_size(4) == 25 and H.sh_size(6) == 49
    assert H.sh_index(0, 0) == 0 and H.sh_index(1, -1) == 1 and H.sh_index(1, 1) == 3 and H.sh_index(2, -2) == 4
    with pytest.raises(ValueError):
        H.sh_index(2, 3)


def test_factorial_table_is_lgamma_close():
    # _sh_indexing.py:120-146: exp(lgamma(n+1)), 1.4e-14 worst relative error up to 20!
    f = H.factorial_table(20)
    for n in range(21):
        assert abs(f[n] - math.factorial(n)) <= 2e-14 * math.factorial(n)


def test_dehnen_table3_degree_one():
    # tests/unit/operators/test_real_harmonics.py:91-118: U_1^-1 = y/2, U_1^0 = z, U_1^1 = x/2
    for d, idx, val in (((1.0, 0, 0), H.sh_index(1, 1), 0.5), ((0, 1.0, 0), H.sh_index(1, -1), 0.5),
                        ((0, 0, 1.0), H.sh_index(1, 0), 1.0)):
        u = H.solid_harmonics(np.array(d, float), 1)
        assert abs(u[0] - 1.0) < 1e-15
        for k in range(1, 4):
            assert abs(u[k] - (val if k == idx else 0.0)) < 1e-12


def test_degree_two_closed_forms():
    # R_2^0 = (2z^2 - x^2 - y^2)/4, R_2^1 = z (x+iy)/2, R_2^2 = (x+iy)^2/8
    x, y, z = 0.3, -0.7, 1.1
    u = H.solid_harmonics(np.array([x, y, z]), 2)
    assert u[H.sh_index(2, 0)] == pytest.approx((2 * z * z - x * x - y * y) / 4, rel=1e-14)
    assert u[H.sh_index(2, 1)] == pytest.approx(z * x / 2, rel=1e-14)
    assert u[H.sh_index(2, -1)] == pytest.approx(z * y / 2, rel=1e-14)
    assert u[H.sh_index(2, 2)] == pytest.approx((x * x - y * y) / 8, rel=1e-14)
    assert u[H.sh_index(2, -2)] == pytest.approx(x * y / 4, rel=1e-14)


def test_z_derivative_lowers_degree():
    # test_real_harmonics.py:150-200: dU_n^m/dz = U_{n-1}^m exactly, 0 when |m| > n-1
    p = 6
    for _ in range(4):
        d = RNG.normal(size=3)
        u, g = H.solid_harmonics_with_grad(d, p)
        for n in range(1, p + 1):
            for m in range(-n, n + 1):
                want = u[H.sh_index(n - 1, m)] if abs(m) <= n - 1 else 0.0
                assert abs(g[H.sh_index(n, m), 2] - want) < 1e-13 * max(1.0, abs(want))


def test_transverse_derivatives_have_half_integer_coefficients():
    # test_real_harmonics.py:200-248: dU/dx, dU/dy are fixed half-integer combos of degree n-1
    p = 6
    fit = [H.solid_harmonics_with_grad(RNG.normal(size=3), p) for _ in range(14)]
    check = [H.solid_harmonics_with_grad(RNG.normal(size=3), p) for _ in range(6)]
    for n in range(1, p + 1):
        lower = list(range(-(n - 1), n))
        fb = np.array([[u[H.sh_index(n - 1, k)] for k in lower] for u, _ in fit])
        cb = np.array([[u[H.sh_index(n - 1, k)] for k in lower] for u, _ in check])
        for m in range(-n, n + 1):
            for axis in (0, 1):
                tgt = np.array([g[H.sh_index(n, m), axis] for _, g in fit])
                sol, *_ = np.linalg.lstsq(fb, tgt, rcond=None)
                held = np.array([g[H.sh_index(n, m), axis] for _, g in check])
                assert np.max(np.abs(cb @ sol - held)) < 1e-11
                assert np.all(np.abs(2 * sol - np.round(2 * sol)) < 1e-9)


def test_gradient_matches_finite_differences():
    p = 5
    d = np.array([0.4, -0.3, 0.8])
    _, g = H.solid_harmonics_with_grad(d, p)
    h = 1e-6
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        fd = (H.solid_harmonics(d + e, p) - H.solid_harmonics(d - e, p)) / (2 * h)
        assert np.max(np.abs(fd - g[:, a])) < 1e-8


def test_p2m_monopole_linearity_and_origin():
    # test_real_harmonics.py:249-285
    d = np.array([[1.0, 2.0, 3.0]])
    for p in (0, 1, 2, 4):
        assert H.p2m(d, np.array([2.5]), p)[0] == pytest.approx(2.5)
    np.testing.assert_allclose(H.p2m(d, np.array([2.0]), 3), 2.0 * H.p2m(d, np.array([1.0]), 3))
    m0 = H.p2m(np.zeros((1, 3)), np.array([1.0]), 4)
    assert m0[0] == 1.0 and np.all(np.abs(m0[1:]) < 1e-10)


def test_l2p_monopole_is_constant_potential():
    F = np.zeros(H.sh_size(4))
    F[0] = 1.0
    for d in ([1.0, 0, 0], [0.5, 0.5, 0.5], [0, 0, 1.0]):
        phi, grad = H.l2p(F, np.array(d), 4)
        assert phi == pytest.approx(1.0) and np.all(grad == 0)


def test_B_matrices_are_involutions():
    # test_real_harmonics.py:426-436
    for ell in range(7):
        BT, BU = H.real_B_matrices(ell)
        eye = np.eye(2 * ell + 1)
        assert np.abs(BU @ BU - eye).max() < 1e-12
        assert np.abs(BT @ BT - eye).max() < 1e-12
        assert np.abs(np.imag(H.dehnen_Q(ell) @ H.dehnen_B_complex(ell).T @ np.linalg.inv(H.dehnen_Q(ell)))).max() < 1e-12


def test_Dz_identity_and_group_law():
    # test_real_harmonics.py:451
    for ell in range(5):
        assert np.allclose(H.Dz(ell, 0.0), np.eye(2 * ell + 1))
        assert np.allclose(H.Dz(ell, 0.3) @ H.Dz(ell, 0.5), H.Dz(ell, 0.8), atol=1e-14)


def _rotation_taking_to_z(d):
    """Physical rotation g with g d/|d| = +z: Rx(ax) Rz(az), az = atan2(x, y), ax = atan2(rho, z)."""
    az, ax = H.alignment_angles(d[0], d[1], d[2])
    Rz = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    return Rx @ Rz


def test_rotation_blocks_match_p2m_of_rotated_source():
    # test_real_harmonics.py:1362: D_to p2m(s) == p2m(g s)
    p = 6
    for _ in range(3):
        d = RNG.normal(size=3)
        g = _rotation_taking_to_z(d)
        assert np.allclose(g @ (d / np.linalg.norm(d)), [0, 0, 1], atol=1e-14)
        s = RNG.normal(size=3)
        lhs = np.zeros(H.sh_size(p))
        u = H.solid_harmonics(s, p)
        for ell in range(p + 1):
            sl = slice(ell * ell, (ell + 1) ** 2)
            lhs[sl] = H.rot_to_z_multipole(d, ell) @ u[sl]
        np.testing.assert_allclose(lhs, H.solid_harmonics(g @ s, p), atol=1e-13)
        for ell in range(p + 1):
            prod = H.rot_from_z_multipole(d, ell) @ H.rot_to_z_multipole(d, ell)
            assert np.abs(prod - np.eye(2 * ell + 1)).max() < 1e-12


def test_z_aligned_rotation_is_identity():
    for ell in range(5):
        assert np.allclose(H.rot_to_z_multipole(np.array([0, 0, 2.0]), ell), np.eye(2 * ell + 1), atol=1e-14)


def test_local_rotation_leaves_potential_invariant():
    # test_real_harmonics.py:1432
    p = 5
    F = RNG.normal(size=H.sh_size(p))
    d = RNG.normal(size=3)
    delta = RNG.normal(size=3) * 0.3
    g = _rotation_taking_to_z(d)
    Fr = np.zeros_like(F)
    for ell in range(p + 1):
        sl = slice(ell * ell, (ell + 1) ** 2)
        Fr[sl] = H.rot_to_z_local(d, ell) @ F[sl]
    assert H.l2p(Fr, g @ delta, p)[0] == pytest.approx(H.l2p(F, delta, p)[0], rel=1e-12)


def test_z_translations_identity_at_zero():
    # test_real_harmonics.py:500-516
    c = RNG.normal(size=H.sh_size(4))
    np.testing.assert_allclose(H.zshift(c, 0.0, 4, "m2m"), c)
    np.testing.assert_allclose(H.zshift(c, 0.0, 4, "l2l"), c)


def test_m2m_is_exact_for_kept_degrees():
    # M2M re-centres a multipole without truncation: m2m(p2m(x - c1), c1 - c0) == p2m(x - c0)
    p = 6
    xs = RNG.normal(size=(5, 3)) * 0.2
    ms = RNG.uniform(0.5, 1.5, size=5)
    c1 = np.array([0.05, -0.02, 0.1])
    c0 = np.array([-0.1, 0.07, 0.0])
    got = H.m2m(H.p2m(xs - c1, ms, p), c1 - c0, p)
    np.testing.assert_allclose(got, H.p2m(xs - c0, ms, p), atol=1e-14)
    # wrong sign of delta is far off
    bad = H.m2m(H.p2m(xs - c1, ms, p), c0 - c1, p)
    assert np.abs(bad - H.p2m(xs - c0, ms, p)).max() > 1e-4


def test_l2l_is_exact_for_polynomial_locals():
    # a local expansion is a degree-p polynomial; l2l re-expands it exactly
    p = 6
    F = RNG.normal(size=H.sh_size(p))
    cp = np.array([0.3, -0.2, 0.5])
    cc = np.array([0.35, -0.1, 0.42])
    x = np.array([0.33, -0.15, 0.47])
    Fc = H.l2l(F, cp - cc, p)
    assert H.l2p(Fc, cc - x, p)[0] == pytest.approx(H.l2p(F, cp - x, p)[0], rel=1e-12)


def test_m2l_monopole_gives_inverse_distance():
    # test_real_harmonics.py:518-538
    M = np.zeros(H.sh_size(4))
    M[0] = 1.0
    assert H.zm2l(M, 2.0, 4)[0] == pytest.approx(0.5, abs=1e-12)
    assert H.m2l(M, np.array([2.0, 0, 0]), 4)[0] == pytest.approx(0.5, abs=1e-12)


def test_z_m2l_converges_with_order():
    # test_real_harmonics.py:541-598
    R, off = 10.0, 0.5
    ev = np.array([off * 0.6, off * 0.8, R - off * 0.2])
    direct = 1.0 / np.linalg.norm(ev)
    errs = []
    for p in (1, 5, 9):
        M = np.zeros(H.sh_size(p))
        M[0] = 1.0
        L = H.zm2l(M, R, p)
        errs.append(abs(H.l2p(L, np.array([0, 0, R]) - ev, p)[0] - direct) / direct)
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-9


# Convention fixture (tests/unit/operators/test_convention_contracts.py:39-54, order 6, float64)
SOURCE_CENTER = np.array([0.10, -0.05, 0.02])
PARENT_CENTER = np.array([0.0, 0.0, 0.0])
PARTICLE = np.array([0.13, -0.02, 0.06])
MASS = 0.83
TARGET_CENTER = np.array([4.0, 1.5, -2.5])
EVAL_POINT = np.array([4.03, 1.47, -2.52])


def _fixture_error(L, centre):
    exact = MASS / np.linalg.norm(EVAL_POINT - PARTICLE)
    return abs(H.l2p(L, centre - EVAL_POINT, 6)[0] - exact) / exact


def test_translation_conventions_fixture():
    # right convention < 1e-9, flipped > 1e-4 (test_convention_contracts.py:152-229)
    M = H.p2m((PARTICLE - SOURCE_CENTER)[None], np.array([MASS]), 6)
    assert _fixture_error(H.m2l(M, TARGET_CENTER - SOURCE_CENTER, 6), TARGET_CENTER) < 1e-9
    assert _fixture_error(H.m2l(M, SOURCE_CENTER - TARGET_CENTER, 6), TARGET_CENTER) > 1e-4
    Mp = H.m2m(M, SOURCE_CENTER - PARENT_CENTER, 6)
    assert _fixture_error(H.m2l(Mp, TARGET_CENTER - PARENT_CENTER, 6), TARGET_CENTER) < 1e-9
    Mw = H.m2m(M, PARENT_CENTER - SOURCE_CENTER, 6)
    assert _fixture_error(H.m2l(Mw, TARGET_CENTER - PARENT_CENTER, 6), TARGET_CENTER) > 1e-4
    child = TARGET_CENTER + np.array([0.02, -0.03, -0.01])
    L = H.m2l(M, TARGET_CENTER - SOURCE_CENTER, 6)
    assert _fixture_error(H.l2l(L, TARGET_CENTER - child, 6), child) < 1e-9
    assert _fixture_error(H.l2l(L, child - TARGET_CENTER, 6), child) > 1e-4
    # L2P offset sign: centre - eval is right, the opposite is ~1.5e-2 off
    exact = MASS / np.linalg.norm(EVAL_POINT - PARTICLE)
    assert abs(H.l2p(L, EVAL_POINT - TARGET_CENTER, 6)[0] - exact) / exact > 1e-3


# docs/operator_conventions.md:138-148: unit mass at (0.2,-0.1,0.3) from the source
# centre, evaluated at (0.15,0.2,-0.1) from the target centre; relative potential error
# of m2l_real at orders 2/4/6 (two significant figures as published).
PER_DIRECTION_TABLE = [
    ((0, 0, 4), (5.2e-05, 1.3e-05, 9.8e-08)),
    ((4, 1.5, -2.5), (4.5e-04, 6.1e-07, 2.7e-08)),
    ((5, 0, 0), (1.1e-04, 1.5e-06, 1.8e-08)),
    ((-2, -3, -1.5), (2.2e-04, 5.2e-06, 1.0e-07)),
    ((1e-6, 0, 5), (3.5e-05, 4.2e-06, 2.1e-08)),
]


@pytest.mark.parametrize("delta,expected", PER_DIRECTION_TABLE)
def test_m2l_per_direction_error_table(delta, expected):
    src = np.array([0.2, -0.1, 0.3])
    ev = np.array([0.15, 0.2, -0.1])
    d = np.array(delta, float)
    for p, want in zip((2, 4, 6), expected):
        L = H.m2l(H.p2m(src[None], np.array([1.0]), p), d, p)
        exact = 1.0 / np.linalg.norm(d + ev - src)
        err = abs(H.l2p(L, -ev, p)[0] - exact) / exact
        assert float(f"{err:.1e}") == want, (delta, p, err, want)


def test_m2l_flop_count():
    assert [H.m2l_flops(p) for p in (2, 4, 6, 8)] == [156, 718, 1964, 4166]
