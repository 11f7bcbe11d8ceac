"""CPU pins of the mutual-lane restatement (oracle/mutual.py).

Mirrors the reference's mutual FMM tests (tests/integration/test_mutual_fmm.py:
129-171 partition and canonical order, :179-193 near-only case, :222-237
momentum, :308-324 level partition, :364-386 uniform-rung level split) on the
float64 oracle, plus two pins of this repository's topology: the reference's
symmetric dual walk yields the same unordered pairs as the oracle topology's
lists (of which the product's C++ builder is a bit-exact copy), and with unit
weights the mutual assembly equals the plain FMM on the same topology.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fmm as ofmm
from oracle import mutual as om
from oracle.topology import build_topology

SOFT = 1e-2


def _system(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 1.0, (n, 3)), rng.uniform(0.5, 1.5, n)


@pytest.mark.parametrize("n, leaf, theta", [(256, 8, 0.6), (512, 16, 0.9), (600, 16, 1.2)])
def test_dual_walk_covers_every_pair_exactly_once(n, leaf, theta):
    x, m = _system(n)
    topo = build_topology(x, m, leaf, theta)
    fa, fb, na, nb = om.dual_traverse(topo, theta)
    assert np.all(fa < fb) and np.all(na < nb)
    count = np.zeros((n, n), np.int32)

    def members(u):
        a, b = topo.node_range[u]
        return np.arange(a, b + 1)

    for a, b in list(zip(fa, fb)) + list(zip(na, nb)):
        ia, ib = members(a), members(b)
        count[np.ix_(ia, ib)] += 1
        count[np.ix_(ib, ia)] += 1
    for l in range(topo.num_leaves):
        il = members(topo.num_internal + l)
        count[np.ix_(il, il)] += 1
    off = count[~np.eye(n, dtype=bool)]
    assert off.min() == 1 and off.max() == 1


@pytest.mark.parametrize("n, leaf, theta, seed", [(700, 8, 0.6, 1), (1024, 16, 0.9, 2), (900, 4, 1.1, 3)])
def test_dual_walk_equals_the_symmetric_topology_lists(n, leaf, theta, seed):
    """The canonical (a < b) halves of the topology's symmetric far and near lists
    are exactly the reference's mutual walk on the same tree."""
    x, m = _system(n, seed)
    topo = build_topology(x, m, leaf, theta)
    fa, fb, na, nb = om.dual_traverse(topo, theta)
    tgt = np.repeat(np.arange(topo.num_nodes), np.diff(topo.far_off))
    src = topo.far_src
    keep = tgt < src
    got_far = set(zip(tgt[keep].tolist(), src[keep].tolist()))
    lt = np.repeat(np.arange(topo.num_leaves), np.diff(topo.near_off))
    ls = topo.near_src
    keep = lt < ls
    I = topo.num_internal
    got_near = set(zip((lt[keep] + I).tolist(), (ls[keep] + I).tolist()))
    assert len(got_far) > 0
    assert got_far == set(zip(fa.tolist(), fb.tolist()))
    assert got_near == set(zip(na.tolist(), nb.tolist()))
    # and the lists are symmetric: every pair appears in both directions
    assert 2 * len(got_far) == topo.far_src.shape[0]


def test_near_only_matches_direct_sum_to_round_off():
    x, m = _system(512)
    acc = om.mutual_accelerations(x, m, leaf_size=32, p=4, theta=0.3, eps=SOFT)
    topo = build_topology(x, m, 32, 0.3)
    assert om.dual_traverse(topo, 0.3)[0].size == 0
    ref = om.level_direct_sum(x, m, eps=SOFT)
    assert np.linalg.norm(acc - ref) / np.linalg.norm(ref) < 1e-13
    assert om.momentum_residual(acc, m) < 1e-14


@pytest.mark.parametrize("theta, p", [(0.6, 2), (0.9, 4), (1.1, 6)])
def test_unit_weights_equal_the_plain_fmm_and_conserve_momentum(theta, p):
    x, m = _system(800, 4)
    topo = build_topology(x, m, 16, theta)
    acc, near, far = om.mutual_accelerations(x, m, leaf_size=16, p=p, theta=theta, eps=SOFT, topo=topo,
                                             return_parts=True)
    plain = ofmm.fmm_accelerations(x, m, leaf_size=16, p=p, theta=theta, eps=SOFT, topo=topo)
    assert np.linalg.norm(acc - plain) / np.linalg.norm(plain) < 1e-12
    ref = om.level_direct_sum(x, m, eps=SOFT)
    err = np.linalg.norm(acc - ref) / np.linalg.norm(ref)
    mom = om.momentum_residual(acc, m)
    assert err > 1e-8  # a far field is present
    assert mom < 1e-14 and mom < 1e-6 * err
    assert om.momentum_residual(far / m[:, None], m) < 1e-14


def test_levels_partition_and_uniform_rung_split():
    n, kmax = 700, 3
    x, m = _system(n, 5)
    topo = build_topology(x, m, 16, 0.9)
    rung = np.random.default_rng(11).integers(0, kmax + 1, n)
    total = om.mutual_accelerations(x, m, leaf_size=16, p=4, theta=0.9, eps=SOFT, topo=topo)
    levels = []
    for k in range(kmax + 1):
        w = np.zeros(kmax + 1)
        w[k] = 1.0
        a = om.mutual_accelerations(x, m, leaf_size=16, p=4, theta=0.9, eps=SOFT, topo=topo, rung=rung,
                                    level_weights=w)
        assert om.momentum_residual(a, m) < 1e-13
        levels.append(a)
    assert np.linalg.norm(sum(levels) - total) / np.linalg.norm(total) < 1e-13
    # one rung for everyone: the level split is the dense oracle's
    uni = np.full(n, 2)
    for k in range(kmax + 1):
        w = np.zeros(kmax + 1)
        w[k] = 1.0
        a = om.mutual_accelerations(x, m, leaf_size=16, p=4, theta=0.9, eps=SOFT, topo=topo, rung=uni,
                                    level_weights=w)
        if k != 2:
            assert np.abs(a).max() == 0.0
        else:
            ref = om.level_direct_sum(x, m, eps=SOFT, rung=uni, level=2)
            assert np.linalg.norm(a - ref) / np.linalg.norm(ref) < 1e-2
