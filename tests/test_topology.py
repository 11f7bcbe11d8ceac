"""Topology spec: oracle invariants and bit-exact agreement of the native builder (CPU)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import ics
from oracle import topology as OT


def _ancestors(topo, node):
    out = []
    while node >= 0:
        out.append(node)
        node = int(topo.parent[node])
    return out


def _check_completeness(topo):
    """Every target leaf sees each particle exactly once via self + near + far(ancestors)."""
    n = topo.num_particles
    I = topo.num_internal
    for l in range(topo.num_leaves):
        hit = np.zeros(n, np.int64)
        a, b = topo.node_range[I + l]
        hit[a : b + 1] += 1
        for s in topo.near_src[topo.near_off[l] : topo.near_off[l + 1]]:
            sa, sb = topo.node_range[I + s]
            hit[sa : sb + 1] += 1
        for t in _ancestors(topo, I + l):
            for s in topo.far_src[topo.far_off[t] : topo.far_off[t + 1]]:
                sa, sb = topo.node_range[s]
                hit[sa : sb + 1] += 1
        assert np.all(hit == 1), (l, np.unique(hit))


CASES = [(1, 4, 0.6), (2, 1, 0.6), (7, 4, 0.6), (64, 1, 0.5), (256, 8, 0.5), (1000, 16, 0.7), (3000, 32, 0.9)]


@pytest.mark.parametrize("n,leaf,theta", CASES)
def test_invariants(n, leaf, theta):
    x, m = ics.uniform(n, n)
    topo = OT.build_topology(x, m, leaf, theta)
    L, I = topo.num_leaves, topo.num_internal
    assert L == (n + leaf - 1) // leaf and I == L - 1 and topo.num_nodes == 2 * L - 1
    # sorted by (code, index)
    assert np.all(np.diff(topo.codes.astype(np.float64)) >= 0)
    assert np.array_equal(np.sort(topo.perm), np.arange(n))
    # radix tree: each internal node's range is the union of its children's, contiguous
    for k in range(I):
        lc, rc = topo.left[k], topo.right[k]
        assert topo.node_range[lc, 0] == topo.node_range[k, 0]
        assert topo.node_range[rc, 1] == topo.node_range[k, 1]
        assert topo.node_range[lc, 1] + 1 == topo.node_range[rc, 0]
        assert topo.parent[lc] == k and topo.parent[rc] == k
    if I:
        assert topo.parent[0] == -1 and topo.node_range[0, 0] == 0 and topo.node_range[0, 1] == n - 1
    # every accepted far pair satisfies the MAC, in both directions
    tgt = np.repeat(np.arange(topo.num_nodes), np.diff(topo.far_off))
    for t, s in zip(tgt, topo.far_src):
        assert OT.mac_accepts(topo, int(t), int(s), theta) and OT.mac_accepts(topo, int(s), int(t), theta)
    # near pairs are symmetric, exclude self, and fail the MAC
    ntg = np.repeat(np.arange(L), np.diff(topo.near_off))
    pairs = set(zip(ntg.tolist(), topo.near_src.tolist()))
    assert all((b, a) in pairs for a, b in pairs) and all(a != b for a, b in pairs)
    for a, b in pairs:
        assert not OT.mac_accepts(topo, I + a, I + b, theta)
    _check_completeness(topo)
    # radius is the exact bounding radius about the centre of mass
    xs, ms = x[topo.perm], m[topo.perm]
    for k in range(topo.num_nodes):
        a, b = topo.node_range[k]
        c = (ms[a : b + 1, None] * xs[a : b + 1]).sum(0) / ms[a : b + 1].sum()
        assert np.allclose(topo.center[k], c, rtol=1e-13, atol=1e-14)
        assert topo.radius[k] == pytest.approx(np.linalg.norm(xs[a : b + 1] - topo.center[k], axis=1).max(), rel=1e-14)


def test_tiny_tree_brute_force():
    """N=8, leaf 1: the radix split of every node is the highest differing Morton bit."""
    x = np.array([[0.1, 0.1, 0.1], [0.9, 0.1, 0.1], [0.1, 0.9, 0.1], [0.9, 0.9, 0.9], [0.5, 0.5, 0.5],
                  [0.2, 0.8, 0.4], [0.7, 0.3, 0.6], [0.35, 0.65, 0.15]])
    m = np.ones(8)
    topo = OT.build_tree(x, m, 1)
    codes = topo.codes
    for k in range(topo.num_internal):
        a, b = topo.node_range[k]
        split = topo.node_range[topo.left[k], 1]
        hb = (int(codes[a]) ^ int(codes[b])).bit_length() - 1
        assert all((int(codes[i]) >> hb) & 1 == 0 for i in range(a, split + 1))
        assert all((int(codes[i]) >> hb) & 1 == 1 for i in range(split + 1, b + 1))


def test_morton_interleave_known_values():
    x = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    c = OT.morton_codes(x, bounds=(np.zeros(3), np.ones(3)))
    assert int(c[0]) == 0 and int(c[1]) == (1 << 63) - 1
    y = np.array([[0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.0, 0.0, 0.5]])
    c = OT.morton_codes(y, bounds=(np.zeros(3), np.ones(3)))
    assert [int(v) for v in c] == [1 << 62, 1 << 61, 1 << 60]


def test_determinism():
    x, m = ics.plummer(2000, 3)
    a = OT.build_topology(x, m, 16, 0.6)
    b = OT.build_topology(x, m, 16, 0.6)
    assert np.array_equal(a.far_src, b.far_src) and np.array_equal(a.near_src, b.near_src)


NATIVE_CASES = [(1, 4, 0.6, "uniform"), (7, 4, 0.6, "uniform"), (256, 8, 0.5, "uniform"),
                (4096, 16, 0.6, "uniform"), (5000, 32, 0.7, "plummer"), (20000, 64, 0.6, "plummer"),
                (3000, 5, 0.55, "blobs")]


@pytest.mark.parametrize("n,leaf,theta,gen", NATIVE_CASES)
def test_native_topology_is_bit_exact(n, leaf, theta, gen):
    from jaccpot_amd.topology import build_topology

    x, m = getattr(ics, gen)(n, 3)
    o = OT.build_topology(x, m, leaf, theta)
    for threads in (1, 0):
        j = build_topology(x, m, leaf_size=leaf, theta=theta, num_threads=threads)
        for name in ("perm", "parent", "left", "right", "node_range", "depth", "center", "radius", "mass", "far_off",
                     "far_src", "near_off", "near_src"):
            assert np.array_equal(getattr(o, name), getattr(j, name)), name
        assert j.max_depth == int(o.depth.max())


def test_native_topology_with_explicit_bounds_and_duplicates():
    from jaccpot_amd.topology import build_topology

    x, m = ics.uniform(500, 4)
    x[100:150] = x[10]  # coincident particles share a Morton code; index breaks the tie
    bounds = (np.full(3, -2.0), np.full(3, 2.0))
    o = OT.build_topology(x, m, 8, 0.6, bounds=bounds)
    j = build_topology(x, m, leaf_size=8, theta=0.6, bounds=bounds)
    assert np.array_equal(o.perm, j.perm) and np.array_equal(o.far_src, j.far_src)
    assert np.array_equal(o.near_src, j.near_src)


def test_native_topology_rejects_bad_input():
    from jaccpot_amd import _native as nat
    from jaccpot_amd.topology import build_topology

    with pytest.raises(nat.JfmError):
        build_topology(np.array([[0.0, np.nan, 0.0]]), np.ones(1), leaf_size=4, theta=0.6)
    with pytest.raises(nat.JfmError):
        build_topology(np.zeros((4, 3)), np.ones(4), leaf_size=4, theta=-1.0)
    with pytest.raises(ValueError):
        build_topology(np.zeros((4, 2)), np.ones(4), leaf_size=4, theta=0.6)
