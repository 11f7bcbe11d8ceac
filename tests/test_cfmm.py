"""The C++/OpenMP FMM restatement (oracle/cfmm.cpp) against the NumPy oracle.

It is the oracle of the GPU path at sizes the NumPy pipeline cannot reach
and the benchmark's CPU baseline, so it is pinned here: on the same topology
it must reproduce oracle.fmm in float64 to rounding (accelerations and
potentials, every order, per-pair orders), while a change of order moves the
result by orders of magnitude more than that tolerance.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cfmm
from oracle import fmm as ofmm
from oracle import ics
from oracle import topology as OT


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not cfmm.LIBRARY.exists():
        cfmm.build()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [
    # (generator, n, leaf, theta, eps)
    ("uniform", 4096, 16, 0.6, 0.0),
    ("plummer", 3000, 16, 0.6, 1e-3),
    ("blobs", 2000, 64, 0.7, 1e-3),
    ("uniform", 7, 4, 0.6, 0.0),
]


@pytest.mark.parametrize("p", [2, 4, 6, 8])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}{c[1]}_leaf{c[2]}" for c in CASES])
def test_matches_numpy_oracle_fp64(case, p):
    gen, n, leaf, theta, eps = case
    x, m = getattr(ics, gen)(n, 5)
    topo = OT.build_topology(x, m, leaf, theta)
    a_ref, p_ref = ofmm.fmm_accelerations(x, m, leaf_size=leaf, p=p, theta=theta, eps=eps, topo=topo,
                                          return_potential=True)
    a, pot = cfmm.evaluate(topo, x, m, p=p, eps=eps, return_potential=True)
    assert _rel(a, a_ref) < 1e-13
    assert _rel(pot, p_ref) < 1e-13


def test_per_pair_orders_match_numpy_oracle():
    x, m = ics.plummer(3000, 3)
    topo = OT.build_topology(x, m, 16, 0.6)
    orders = np.random.default_rng(0).integers(0, 7, size=topo.far_src.shape[0])
    ref = ofmm.fmm_accelerations(x, m, leaf_size=16, p=6, theta=0.6, eps=1e-3, topo=topo, orders=orders)
    got = cfmm.evaluate(topo, x, m, p=6, eps=1e-3, orders=orders)
    fixed = cfmm.evaluate(topo, x, m, p=6, eps=1e-3)
    assert _rel(got, ref) < 1e-13
    # the far field is actually exercised: lowering orders moves the result far beyond the tolerance
    assert _rel(fixed, ref) > 1e-5


def test_orders_change_the_far_field():
    """Each order matches the oracle to 1e-13 while orders 2 and 8 differ by far more."""
    x, m = ics.uniform(2048, 9)
    topo = OT.build_topology(x, m, 8, 0.6)
    a2 = cfmm.evaluate(topo, x, m, p=2)
    a8 = cfmm.evaluate(topo, x, m, p=8)
    assert _rel(a2, a8) > 1e-5


def test_float32_build_close_to_float64():
    x, m = ics.plummer(20000, 4)
    topo = OT.build_topology(x, m, 32, 0.6)
    a64 = cfmm.evaluate(topo, x, m, p=6, eps=1e-3)
    a32 = cfmm.evaluate(topo, x, m, p=6, eps=1e-3, precision="fp32")
    # float32 arithmetic with plain (uncompensated) float32 sums: measured 2.0e-5
    assert _rel(a32, a64) < 1e-4


def test_native_topology_gives_same_result():
    """The product's host topology builder is bit-identical to the oracle's, so the
    restatement can run on either (the benchmark's CPU baseline uses the native one)."""
    torch = pytest.importorskip("torch")  # noqa: F841  (libjfm shares torch's HIP runtime)
    from jaccpot_amd.topology import build_topology

    x, m = ics.plummer(5000, 6)
    t_or = OT.build_topology(x, m, 16, 0.6)
    t_nat = build_topology(x, m, leaf_size=16, theta=0.6)
    a = cfmm.evaluate(t_or, x, m, p=4, eps=1e-3)
    b = cfmm.evaluate(t_nat, x, m, p=4, eps=1e-3)
    assert np.array_equal(a, b)


def test_rejects_bad_orders():
    x, m = ics.uniform(500, 1)
    topo = OT.build_topology(x, m, 16, 0.6)
    with pytest.raises(ValueError):
        cfmm.evaluate(topo, x, m, p=4, orders=np.full(topo.far_src.shape[0], 5))
    with pytest.raises(ValueError):
        cfmm.evaluate(topo, x, m, p=4, orders=np.zeros(3))
