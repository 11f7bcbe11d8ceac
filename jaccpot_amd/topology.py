"""Fixed static-radix topology, built natively by libjfm (host C++).

Plays the role of yggdrax's tree + interaction-list builders for the reference
(jaccpot/runtime/fmm_state.py:630-800, runtime/_interaction_cache.py:523,564):
Morton sort, leaf buckets, binary radix tree, centres of mass and bounding
radii, and the dual-tree far/near lists for a MAC opening angle theta.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat


@dataclass
class Topology:
    """Host arrays of one topology (int32 indices, float64 geometry)."""

    num_particles: int
    leaf_size: int
    num_leaves: int
    num_internal: int
    num_nodes: int
    max_depth: int
    theta: float
    perm: np.ndarray
    parent: np.ndarray
    left: np.ndarray
    right: np.ndarray
    node_range: np.ndarray
    depth: np.ndarray
    center: np.ndarray
    radius: np.ndarray
    mass: np.ndarray
    far_off: np.ndarray
    far_src: np.ndarray
    near_off: np.ndarray
    near_src: np.ndarray
    m2m_nodes: np.ndarray
    m2m_off: np.ndarray
    l2l_nodes: np.ndarray
    l2l_off: np.ndarray
    m2l_targets: np.ndarray
    build_seconds: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def num_far_pairs(self) -> int:
        return int(self.far_src.shape[0])

    @property
    def num_near_entries(self) -> int:
        return int(self.near_src.shape[0])

    def p2p_interactions(self) -> int:
        """Pairwise interactions the P2P kernel evaluates (self pairs excluded)."""
        counts = (self.node_range[self.num_internal:, 1] - self.node_range[self.num_internal:, 0] + 1).astype(np.int64)
        near_counts = np.zeros(self.num_leaves, np.int64)
        tgt = np.repeat(np.arange(self.num_leaves), np.diff(self.near_off))
        np.add.at(near_counts, tgt, counts[self.near_src])
        return int(np.sum(counts * (near_counts + counts - 1)))


def build_topology(positions, masses, *, leaf_size: int, theta: float, bounds=None, num_threads: int = 0) -> Topology:
    """Build the topology for input-order positions [N,3] / masses [N]."""
    import time

    lib = nat.load()
    x = np.ascontiguousarray(np.asarray(positions, dtype=np.float64))
    m = np.ascontiguousarray(np.asarray(masses, dtype=np.float64))
    if x.ndim != 2 or x.shape[1] != 3:
        raise ValueError("positions must have shape (N, 3)")
    if m.shape != (x.shape[0],):
        raise ValueError("masses must have shape (N,)")
    n = x.shape[0]
    if leaf_size < 1 or leaf_size > 256:
        raise ValueError("leaf_size must be in [1, 256]")
    b = None
    if bounds is not None:
        lo, hi = bounds
        b = np.ascontiguousarray(np.concatenate([np.asarray(lo, np.float64), np.asarray(hi, np.float64)]))
    handle = ctypes.c_void_p()
    t0 = time.perf_counter()
    nat.check(
        lib.jfm_topology_build(nat.ptr(x), nat.ptr(m), n, int(leaf_size), float(theta),
                               nat.ptr(b) if b is not None else None, int(num_threads), ctypes.byref(handle)),
        "jfm_topology_build",
    )
    try:
        info = np.zeros(9, np.int64)
        nat.check(lib.jfm_topology_info(handle, nat.ptr(info)), "jfm_topology_info")
        N, leaf, L, I, nodes, K, E, D, T = (int(v) for v in info)
        perm = np.empty(N, np.int32)
        parent = np.empty(nodes, np.int32)
        left = np.empty(max(I, 0), np.int32)
        right = np.empty(max(I, 0), np.int32)
        node_range = np.empty((nodes, 2), np.int32)
        depth = np.empty(nodes, np.int32)
        center = np.empty((nodes, 3), np.float64)
        radius = np.empty(nodes, np.float64)
        mass = np.empty(nodes, np.float64)
        far_off = np.empty(nodes + 1, np.int32)
        far_src = np.empty(K, np.int32)
        near_off = np.empty(L + 1, np.int32)
        near_src = np.empty(E, np.int32)
        nat.check(
            lib.jfm_topology_export(handle, *(nat.ptr(a) if a.size else None for a in (
                perm, parent, left, right, node_range, depth, center, radius, mass, far_off, far_src, near_off,
                near_src))),
            "jfm_topology_export",
        )
        m2m_nodes = np.empty(max(I, 0), np.int32)
        m2m_off = np.zeros(D + 2, np.int32)
        l2l_nodes = np.empty(max(nodes - 1, 0), np.int32)
        l2l_off = np.zeros(D + 2, np.int32)
        m2l_targets = np.empty(T, np.int32)
        nat.check(
            lib.jfm_topology_schedules(handle, nat.ptr(m2m_nodes) if m2m_nodes.size else None, nat.ptr(m2m_off),
                                       nat.ptr(l2l_nodes) if l2l_nodes.size else None, nat.ptr(l2l_off),
                                       nat.ptr(m2l_targets) if m2l_targets.size else None),
            "jfm_topology_schedules",
        )
    finally:
        lib.jfm_topology_free(handle)
    t1 = time.perf_counter()
    return Topology(
        num_particles=N, leaf_size=leaf, num_leaves=L, num_internal=I, num_nodes=nodes, max_depth=D,
        theta=float(theta), perm=perm, parent=parent, left=left, right=right, node_range=node_range, depth=depth,
        center=center, radius=radius, mass=mass, far_off=far_off, far_src=far_src, near_off=near_off,
        near_src=near_src, m2m_nodes=m2m_nodes, m2m_off=m2m_off[: D + 2] if I > 0 else m2m_off[:1],
        l2l_nodes=l2l_nodes, l2l_off=l2l_off[: D + 1] if D > 0 else l2l_off[:1], m2l_targets=m2l_targets,
        build_seconds=t1 - t0,
    )
