"""Momentum-conserving (mutual) FMM with block-step level weights, float64
NumPy restatement (oracle / test infrastructure only).

Restates the reference's mutual lane:

* ``dual_traverse`` -- the symmetric dual-tree walk of
  jaccpot/mutual/topology.py:300-395: a frontier of unordered node pairs from
  (root, root); a pair a != b is far when ``theta |c_b - c_a| > r_a + r_b``;
  two leaves (a != b) are near; otherwise the larger-radius node is split
  (the other one when the larger is a leaf; a self pair splits into
  (l, l), (r, r), (min, max)).  Output pairs are canonical (a < b).
* ``mutual_forces`` -- force.py:885-1068, nearfield.py, farfield.py on a
  given topology: near pair forces ``G m_i m_j w[max(r_i, r_j)] d / r^3``
  (self leaf with i != j, canonical near leaf pairs applied +f / -f); far
  pairs (both directions of each canonical pair) through M2L scaled by
  ``w[max(R_A, R_B)]`` with the cell rung R the maximum over the cell's
  particles (-1 for none, clipped to [0, levels - 1]), then L2L and L2P; far
  force = m a_far; acceleration = force / m (0 for m = 0).
* ``level_direct_sum`` -- the dense per-level reference the reference's tests
  inline (tests/integration/test_mutual_fmm.py:52-89): pairs with
  ``max(r_i, r_j) == level``, ``c_ij`` symmetric, ``d_ij = -d_ji``.

Pinned by: the partition test (every particle pair covered exactly once by
self leaves, near pairs and far pairs), equality of the canonical lists with
this repository's symmetric topology (topology.py) on seeded inputs, the
near-only case against the direct sum to round-off, unit weights against
``fmm.fmm_accelerations`` on the same topology, and momentum conservation to
round-off (tests/test_oracle_mutual.py).
"""

from __future__ import annotations

import numpy as np

from . import harmonics as H
from .fmm import upward
from .topology import Topology, build_topology, nodes_by_depth_asc


def dual_traverse(topo: Topology, theta: float):
    """(far_a, far_b, near_a, near_b) canonical node pairs (a < b), mutual/topology.py:300-395."""
    I = topo.num_internal
    c, r = topo.center, topo.radius
    far_a, far_b, near_a, near_b = [], [], [], []
    cur_a = np.array([0], np.int64)
    cur_b = np.array([0], np.int64)
    theta = float(theta)
    while cur_a.size:
        a, b = cur_a, cur_b
        delta = c[b] - c[a]
        dist = np.sqrt(np.einsum("ij,ij->i", delta, delta))
        acc = (a != b) & (theta * dist > r[a] + r[b])
        far_a.append(a[acc])
        far_b.append(b[acc])
        a, b = a[~acc], b[~acc]
        if a.size == 0:
            break
        leaves = (a >= I) & (b >= I)
        cross = leaves & (a != b)
        near_a.append(a[cross])
        near_b.append(b[cross])
        a, b = a[~leaves], b[~leaves]
        if a.size == 0:
            break
        nxt_a, nxt_b = [], []
        diag = a == b
        if diag.any():
            u = a[diag]
            lc, rc = topo.left[u], topo.right[u]
            nxt_a += [lc, rc, np.minimum(lc, rc)]
            nxt_b += [lc, rc, np.maximum(lc, rc)]
        off = ~diag
        if off.any():
            oa, ob = a[off], b[off]
            a_leaf, b_leaf = oa >= I, ob >= I
            split_a = (~a_leaf) & (b_leaf | (r[oa] >= r[ob]))
            for par, other, msk in ((oa, ob, split_a), (ob, oa, ~split_a)):
                if not msk.any():
                    continue
                p_, o_ = par[msk], other[msk]
                for ch in (topo.left[p_], topo.right[p_]):
                    nxt_a.append(np.minimum(ch, o_))
                    nxt_b.append(np.maximum(ch, o_))
        cur_a = np.concatenate(nxt_a)
        cur_b = np.concatenate(nxt_b)

    def cat(parts):
        return np.concatenate(parts).astype(np.int64) if parts else np.zeros(0, np.int64)

    return cat(far_a), cat(far_b), cat(near_a), cat(near_b)


def cell_rungs(topo: Topology, rung_sorted: np.ndarray) -> np.ndarray:
    """Rung of each node = max over its particles (force.py:885-917)."""
    out = np.full(topo.num_nodes, -1, np.int64)
    for u in range(topo.num_nodes):
        a, b = topo.node_range[u]
        if b >= a:
            out[u] = int(np.max(rung_sorted[a : b + 1]))
    return out


def mutual_forces(topo: Topology, xs: np.ndarray, ms: np.ndarray, p: int, *, theta: float, G: float, eps: float,
                  rung_sorted=None, level_weights=None):
    """(near_force, far_force) [N, 3] in tree order on ``topo`` (canonical pairs from ``dual_traverse``)."""
    n = xs.shape[0]
    I = topo.num_internal
    eps2 = float(eps) * float(eps)
    weighted = rung_sorted is not None and level_weights is not None
    lw = np.asarray(level_weights, np.float64) if weighted else None
    nl = lw.shape[0] if weighted else 1
    far_a, far_b, near_a, near_b = dual_traverse(topo, theta)

    def block(ia, ib, both):
        """Pair forces of particle blocks ia (receivers of +f) and ib (-f when both)."""
        d = xs[ib][None, :, :] - xs[ia][:, None, :]
        r2 = (d * d).sum(-1) + eps2
        live = ia[:, None] != ib[None, :]
        inv = np.where(live, 1.0 / np.sqrt(np.where(live, r2, 1.0)), 0.0)
        c = ms[ia][:, None] * ms[ib][None, :]
        if weighted:
            lv = np.clip(np.maximum(rung_sorted[ia][:, None], rung_sorted[ib][None, :]), 0, nl - 1)
            c = c * lw[lv]
        f = (c * inv ** 3 * G)[..., None] * d
        near[ia] += f.sum(1)
        if both:
            near[ib] -= f.sum(0)

    near = np.zeros((n, 3))
    for l in range(topo.num_leaves):
        a, b = topo.node_range[I + l]
        idx = np.arange(a, b + 1)
        block(idx, idx, False)
    for u, v in zip(near_a, near_b):
        (a0, a1), (b0, b1) = topo.node_range[u], topo.node_range[v]
        block(np.arange(a0, a1 + 1), np.arange(b0, b1 + 1), True)

    # far field: both directions of every canonical pair, cell-level weights
    C = H.sh_size(p)
    M = upward(topo, xs, ms, p)
    Lc = np.zeros((topo.num_nodes, C))
    cen = topo.center
    if far_a.size:
        tgt = np.concatenate([far_a, far_b])
        src = np.concatenate([far_b, far_a])
        contrib = H.m2l(M[src], cen[tgt] - cen[src], p)
        if weighted:
            nr = cell_rungs(topo, rung_sorted)
            lv = np.clip(np.maximum(nr[tgt], nr[src]), 0, nl - 1)
            contrib = contrib * lw[lv][:, None]
        np.add.at(Lc, tgt, contrib)
    for level in nodes_by_depth_asc(topo):
        par = topo.parent[level]
        Lc[level] = Lc[level] + H.l2l(Lc[par], cen[par] - cen[level], p)
    far = np.zeros((n, 3))
    for l in range(topo.num_leaves):
        node = I + l
        a, b = topo.node_range[node]
        _, grad = H.l2p(Lc[node], cen[node] - xs[a : b + 1], p)
        far[a : b + 1] = ms[a : b + 1, None] * (-G * grad)
    return near, far


def mutual_accelerations(positions, masses, *, leaf_size: int, p: int, theta: float, G: float = 1.0,
                         eps: float = 0.0, rung=None, level_weights=None, topo: Topology | None = None,
                         return_parts: bool = False):
    """Accelerations (and near/far forces) in input order, force.py:985-1068."""
    x = np.asarray(positions, np.float64)
    m = np.asarray(masses, np.float64)
    if topo is None:
        topo = build_topology(x, m, leaf_size, theta)
    xs, ms = x[topo.perm], m[topo.perm]
    rs = None if rung is None else np.asarray(rung, np.int64)[topo.perm]
    near_s, far_s = mutual_forces(topo, xs, ms, p, theta=theta, G=G, eps=eps, rung_sorted=rs,
                                  level_weights=level_weights)
    tot = near_s + far_s
    safe = np.where(ms > 0, ms, 1.0)
    acc_s = np.where(ms[:, None] > 0, tot / safe[:, None], 0.0)
    acc = np.empty_like(acc_s)
    near = np.empty_like(acc_s)
    far = np.empty_like(acc_s)
    acc[topo.perm], near[topo.perm], far[topo.perm] = acc_s, near_s, far_s
    if return_parts:
        return acc, near, far
    return acc


def level_direct_sum(positions, masses, *, eps: float, G: float = 1.0, rung=None, level=None):
    """Dense momentum-conserving direct sum, restricted to pairs with max(rung) == level when given."""
    x = np.asarray(positions, np.float64)
    m = np.asarray(masses, np.float64)
    d = x[None, :, :] - x[:, None, :]
    r2 = (d * d).sum(-1) + float(eps) ** 2
    live = ~np.eye(x.shape[0], dtype=bool)
    if rung is not None and level is not None:
        r = np.asarray(rung)
        live &= np.maximum(r[:, None], r[None, :]) == int(level)
    inv3 = np.where(live, np.where(live, r2, 1.0) ** -1.5, 0.0)
    c = G * m[:, None] * m[None, :] * inv3
    return (c[..., None] * d).sum(1) / m[:, None]


def momentum_residual(acc, masses) -> float:
    """|sum_i m_i a_i| over the norm of the per-component sums of |m_i a_i|."""
    terms = np.asarray(masses, np.float64)[:, None] * np.asarray(acc, np.float64)
    scale = np.abs(terms).sum(0)
    return float(np.linalg.norm(terms.sum(0)) / (np.linalg.norm(scale) + 1e-300))
