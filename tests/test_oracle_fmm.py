"""Oracle FMM and direct sum against physics (CPU).

Anchors from the reference: FMM vs direct sum rel-L2 < 1e-2 on uniform N=256,
theta 0.5, leaf 8, eps 1e-2, with measured 7.2e-4 / 8.15e-5 / 1.13e-5 at
p = 2 / 4 / 6 (tests/characterization/test_fmm_golden.py:67-74, :413-416);
convergence in p; root monopole equals the total mass
(tests/unit/runtime/test_upward_node_mass_conservation.py).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fmm as ofmm
from oracle import harmonics as H
from oracle import ics
from oracle import topology as OT
from oracle.direct import direct_sum, relative_errors


def test_direct_sum_two_body_closed_form():
    x = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    m = np.array([3.0, 5.0])
    a, phi = direct_sum(x, m, G=2.0, eps=0.0, return_potential=True)
    np.testing.assert_allclose(a[0], [2.0 * 5.0 / 4.0, 0, 0])
    np.testing.assert_allclose(a[1], [-2.0 * 3.0 / 4.0, 0, 0])
    np.testing.assert_allclose(phi, [-2.0 * 5.0 / 2.0, -2.0 * 3.0 / 2.0])
    a_s = direct_sum(x, m, eps=1.0)
    np.testing.assert_allclose(a_s[0], [5.0 * 2.0 / 5.0 ** 1.5, 0, 0])


def test_direct_sum_newton_third_law():
    x, m = ics.blobs(300, 1)
    a = direct_sum(x, m, eps=1e-3)
    assert np.linalg.norm((m[:, None] * a).sum(0)) < 1e-12 * np.abs(m[:, None] * a).sum()


@pytest.mark.parametrize("p,bound", [(2, 1.5e-3), (4, 1.5e-4), (6, 2.5e-5)])
def test_physics_anchor_uniform256(p, bound):
    rng = np.random.default_rng(0xC0FFEE)
    x = rng.uniform(-1, 1, (256, 3))
    m = np.abs(rng.normal(size=256)) + 0.5
    a = ofmm.fmm_accelerations(x, m, leaf_size=8, p=p, theta=0.5, eps=1e-2)
    ad = direct_sum(x, m, eps=1e-2)
    rel = np.linalg.norm(a - ad) / np.linalg.norm(ad)
    assert rel < 1e-2 and rel < bound, rel


def test_convergence_in_order():
    x, m = ics.plummer(3000, 2)
    ad = direct_sum(x, m, eps=1e-3)
    topo = OT.build_topology(x, m, 16, 0.6)
    errs = [np.median(relative_errors(ofmm.fmm_accelerations(x, m, leaf_size=16, p=p, theta=0.6, eps=1e-3,
                                                             topo=topo), ad)) for p in (2, 4, 6, 8)]
    assert errs[0] > errs[1] > errs[2] > errs[3]
    assert errs[3] < 1e-6


def test_root_monopole_is_total_mass_and_potential_consistent():
    x, m = ics.uniform(2000, 5)
    topo = OT.build_topology(x, m, 16, 0.6)
    M = ofmm.upward(topo, x[topo.perm], m[topo.perm], 4)
    assert M[0, 0] == pytest.approx(m.sum(), rel=1e-13)
    # dipole about the COM vanishes at the root
    assert np.abs(M[0, 1:4]).max() < 1e-12 * m.sum()
    a, phi = ofmm.fmm_accelerations(x, m, leaf_size=16, p=6, theta=0.5, eps=1e-2, topo=topo, return_potential=True)
    ad, pd = direct_sum(x, m, eps=1e-2, return_potential=True)
    assert np.linalg.norm(phi - pd) / np.linalg.norm(pd) < 1e-5


def test_per_pair_orders_reduce_to_fixed_order():
    x, m = ics.uniform(1500, 6)
    topo = OT.build_topology(x, m, 16, 0.6)
    K = topo.far_src.size
    a4 = ofmm.fmm_accelerations(x, m, leaf_size=16, p=4, theta=0.6, topo=topo)
    a4g = ofmm.fmm_accelerations(x, m, leaf_size=16, p=4, theta=0.6, topo=topo, orders=np.full(K, 4))
    np.testing.assert_allclose(a4g, a4, rtol=1e-13, atol=1e-13)
    # a pair evaluated at gear g equals the order-g translation of the truncated multipole
    M = ofmm.upward(topo, x[topo.perm], m[topo.perm], 6)
    k = K // 2
    t = int(np.searchsorted(topo.far_off, k, side="right") - 1)
    s = int(topo.far_src[k])
    d = topo.center[t] - topo.center[s]
    lo = H.m2l(H.truncate(M[s], 6, 2), d, 2)
    hi = H.m2l(M[s], d, 6)
    assert np.abs(lo[0] - hi[0]) < 0.05 * abs(hi[0])


def test_direct_sum_vjp_matches_finite_differences():
    from oracle.direct import direct_sum_vjp

    rng = np.random.default_rng(4)
    x = rng.normal(size=(40, 3))
    m = rng.uniform(0.5, 1.5, 40)
    ab = rng.normal(size=(40, 3))
    pb = rng.normal(size=40)
    for eps in (0.0, 0.05):
        gx, gm = direct_sum_vjp(x, m, ab, pb, G=1.3, eps=eps)

        def loss(xx, mm):
            a, phi = direct_sum(xx, mm, G=1.3, eps=eps, return_potential=True)
            return float((a * ab).sum() + (phi * pb).sum())

        for _ in range(5):
            vx = rng.normal(size=x.shape)
            vm = rng.normal(size=m.shape)
            h = 1e-6
            fd = (loss(x + h * vx, m + h * vm) - loss(x - h * vx, m - h * vm)) / (2 * h)
            an = float((gx * vx).sum() + (gm * vm).sum())
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
