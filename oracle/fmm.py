"""Full fixed-order FMM on the real basis, float64 NumPy restatement.

Oracle / test infrastructure only (see ``oracle/__init__.py``).

Pipeline (runtime/fmm_prepare.py:714-1678, runtime/kernels/_evaluate.py:769-1170):
  P2M over leaves about their centres of mass (upward/real_tree_expansions.py:89-213);
  M2M deepest -> root, parent = m2m(left) + m2m(right) (:216-364);
  M2L over the far list, summed per target (runtime/kernels/_m2l.py:1483-1664);
  L2L root -> leaves, child += l2l(parent) (runtime/kernels/_l2l.py:210-409);
  L2P: a_far = -G grad_delta phi, pot_far = -G phi, delta = centre - x
       (_evaluate.py:1143-1170, :2106-2485);
  P2P: self leaf (i != j) + near leaves, softened Newton (nearfield/_kernels.py:40-338);
  a = a_near + a_far, returned in input order (fmm_evaluate.py:464-540).

Pinned by: the operator pins in ``harmonics``, the topology invariants, and
end-to-end against ``direct.direct_sum`` (physics anchor rel-L2 < 1e-2 and
convergence in p, tests/characterization/test_fmm_golden.py:196-269).
"""

from __future__ import annotations

import numpy as np

from . import harmonics as H
from .topology import Topology, build_topology, internal_nodes_by_depth_desc, nodes_by_depth_asc


def upward(topo: Topology, xs: np.ndarray, ms: np.ndarray, p: int):
    """Leaf P2M then M2M by level; returns multipoles [nodes, C] (centres from topo)."""
    C = H.sh_size(p)
    I = topo.num_internal
    M = np.zeros((topo.num_nodes, C))
    c = topo.center
    for l in range(topo.num_leaves):
        node = I + l
        a, b = topo.node_range[node]
        M[node] = H.p2m(xs[a : b + 1] - c[node], ms[a : b + 1], p)
    for level in internal_nodes_by_depth_desc(topo):
        lc = topo.left[level]
        rc = topo.right[level]
        ml = H.m2m(M[lc], c[lc] - c[level], p)
        mr = H.m2m(M[rc], c[rc] - c[level], p)
        M[level] = ml + mr
    return M


def downward(topo: Topology, M: np.ndarray, p: int, orders=None):
    """M2L over the far list (optionally per-pair order), then the L2L cascade.

    ``orders`` (optional, [K]) gives each far pair's expansion order g <= p: the
    source multipole is truncated to degree g, translated at order g, and added
    into the degrees <= g of the target local.
    """
    C = H.sh_size(p)
    Lc = np.zeros((topo.num_nodes, C))
    c = topo.center
    K = topo.far_src.shape[0]
    if K:
        tgt = np.repeat(np.arange(topo.num_nodes), np.diff(topo.far_off))
        src = topo.far_src
        delta = c[tgt] - c[src]
        if orders is None:
            contrib = H.m2l(M[src], delta, p)
        else:
            contrib = np.zeros((K, C))
            for g in np.unique(orders):
                sel = orders == g
                cg = H.sh_size(int(g))
                contrib[sel, :cg] = H.m2l(M[src[sel], :cg], delta[sel], int(g))
        # per-target sums over each target's contiguous, source-ascending segment
        nz = np.nonzero(np.diff(topo.far_off))[0]
        Lc[nz] = np.add.reduceat(contrib, topo.far_off[nz], axis=0)
    for level in nodes_by_depth_asc(topo):
        par = topo.parent[level]
        Lc[level] = Lc[level] + H.l2l(Lc[par], c[par] - c[level], p)
    return Lc


def evaluate(topo: Topology, xs: np.ndarray, ms: np.ndarray, Lc: np.ndarray, p: int, G: float, eps: float):
    """L2P + P2P in sorted order; returns (acc [N,3], pot [N])."""
    n = xs.shape[0]
    I = topo.num_internal
    acc = np.zeros((n, 3))
    pot = np.zeros(n)
    eps2 = float(eps) * float(eps)
    c = topo.center
    for l in range(topo.num_leaves):
        node = I + l
        a, b = topo.node_range[node]
        xt = xs[a : b + 1]
        # far field
        phi, grad = H.l2p(Lc[node], c[node] - xt, p)
        far_acc = -G * grad
        far_pot = -G * phi
        # near field: self leaf with i != j, then neighbour leaves
        srcs = [np.arange(a, b + 1)]
        for s in topo.near_src[topo.near_off[l] : topo.near_off[l + 1]]:
            sa, sb = topo.node_range[I + s]
            srcs.append(np.arange(sa, sb + 1))
        sidx = np.concatenate(srcs)
        d = xt[:, None, :] - xs[sidx][None, :, :]
        r2 = (d * d).sum(-1) + eps2
        mask = np.ones_like(r2, dtype=bool)
        cnt = b - a + 1
        mask[np.arange(cnt), np.arange(cnt)] = False
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = np.where(mask, 1.0 / np.sqrt(np.where(mask, r2, 1.0)), 0.0)
        w = ms[sidx][None, :] * inv
        near_pot = -G * w.sum(axis=1)
        near_acc = -G * ((w * inv * inv)[..., None] * d).sum(axis=1)
        acc[a : b + 1] = near_acc + far_acc
        pot[a : b + 1] = near_pot + far_pot
    return acc, pot


def fmm_accelerations(positions, masses, *, leaf_size: int, p: int, theta: float, G: float = 1.0,
                      eps: float = 0.0, bounds=None, topo: Topology | None = None, orders=None,
                      return_potential: bool = False):
    """Full FMM in float64; accelerations (and potentials) in input order."""
    x = np.asarray(positions, np.float64)
    m = np.asarray(masses, np.float64)
    if topo is None:
        topo = build_topology(x, m, leaf_size, theta, bounds)
    xs = x[topo.perm]
    ms = m[topo.perm]
    M = upward(topo, xs, ms, p)
    Lc = downward(topo, M, p, orders)
    acc_s, pot_s = evaluate(topo, xs, ms, Lc, p, G, eps)
    acc = np.empty_like(acc_s)
    pot = np.empty_like(pot_s)
    acc[topo.perm] = acc_s
    pot[topo.perm] = pot_s
    if return_potential:
        return acc, pot
    return acc
