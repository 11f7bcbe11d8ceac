"""Fixed static-radix topology spec, float64 NumPy / pure-Python restatement.

Oracle / test infrastructure only (see ``oracle/__init__.py``).

The reference obtains its tree and interaction lists from yggdrax, which is not
vendored (pyproject.toml:89) and cannot be installed offline, so bit parity
with yggdrax's lists is unpinnable here.  This module states the topology that
this repository uses instead, following the structure the reference fixes
around those calls:

* Morton sort with index tie-break (runtime/fmm_prepare.py:380-383);
* leaves = ceil(N / leaf_size) consecutive buckets, nodes = 2L - 1, internal
  nodes [0, L-1), leaves [L-1, 2L-1) (fmm_prepare.py:384-397,
  upward/real_tree_expansions.py:127-128);
* binary radix tree over the leaves' Morton keys, root = node 0;
* centres of mass per node, falling back to the plain mean for a massless node,
  and the exact max-COM-distance bounding radius (mutual/device_topology.py:50-160);
* dual-tree walk with the sum-form MAC (r_t + r_s)^2 <= theta^2 d^2
  (runtime/_interaction_cache.py:1252-1258, experimental/treecode_walk.py:48-66),
  splitting the larger-radius node (mutual/topology.py:300-395), emitting both
  directions of every accepted pair (mutual/topology.py:5-8); the self leaf is
  not in the near list (pallas/nearfield_fused_leaf.py:16-19).

The C++ builder in ``jaccpot_amd/csrc/topology.cpp`` implements the same spec
independently (parallel Karras construction instead of this top-down
recursion); tests require the two to agree bit-exactly.

Parity pins (tests/test_oracle_topology.py): partition completeness (every
target leaf sees every particle exactly once through self + near + far), every
accepted far pair satisfies the MAC, brute-force enumeration of the tree on tiny
N, determinism.
"""

from __future__ import annotations

import sys
from dataclasses import dataclass

import numpy as np

MORTON_BITS = 21


def _spread3(v: np.ndarray) -> np.ndarray:
    """Interleave the low 21 bits of v with two zero bits between each."""
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    out = np.zeros_like(v)
    for b in range(MORTON_BITS):
        out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return out


def morton_codes(positions: np.ndarray, bounds=None) -> np.ndarray:
    """63-bit Morton codes on the cube spanned by ``bounds`` (default: bounding box).

    q_a = floor((x_a - lo_a) / extent * 2^21) clipped to [0, 2^21 - 1], with
    extent = max_a(hi_a - lo_a) (1 if zero); bit 3b+2 is x, 3b+1 is y, 3b is z.
    """
    x = np.asarray(positions, np.float64)
    if bounds is None:
        lo = x.min(axis=0)
        hi = x.max(axis=0)
    else:
        lo = np.asarray(bounds[0], np.float64)
        hi = np.asarray(bounds[1], np.float64)
    extent = float(np.max(hi - lo))
    if not extent > 0.0:
        extent = 1.0
    scale = float(1 << MORTON_BITS) / extent
    q = np.floor((x - lo) * scale)
    q = np.clip(q, 0, (1 << MORTON_BITS) - 1).astype(np.uint64)
    return (_spread3(q[:, 0]) << np.uint64(2)) | (_spread3(q[:, 1]) << np.uint64(1)) | _spread3(q[:, 2])


@dataclass
class Topology:
    """Host description of the fixed topology (all arrays int64 / float64)."""

    num_particles: int
    leaf_size: int
    num_leaves: int
    num_internal: int
    num_nodes: int
    perm: np.ndarray  # sorted slot -> original index
    codes: np.ndarray  # sorted Morton codes
    parent: np.ndarray  # [nodes], -1 at the root
    left: np.ndarray  # [I]
    right: np.ndarray  # [I]
    node_range: np.ndarray  # [nodes, 2] inclusive sorted-particle range
    depth: np.ndarray  # [nodes]
    mass: np.ndarray  # [nodes]
    center: np.ndarray  # [nodes, 3]
    radius: np.ndarray  # [nodes]
    far_off: np.ndarray  # [nodes + 1] CSR by target node
    far_src: np.ndarray  # [K]
    near_off: np.ndarray  # [L + 1] CSR by target leaf (leaf indices, self excluded)
    near_src: np.ndarray  # [E]

    def leaf_node(self, leaf: int) -> int:
        return self.num_internal + leaf


def _delta(keys: np.ndarray, i: int, j: int) -> int:
    """Common-prefix length of the (code, index) keys of leaves i and j."""
    a = int(keys[i])
    b = int(keys[j])
    if a != b:
        return 64 - (a ^ b).bit_length()
    return 64 + 32 - (i ^ j).bit_length()


def _find_split(keys: np.ndarray, first: int, last: int) -> int:
    """Last index s in [first, last) whose key shares more prefix with key[first]
    than key[last] does (the highest differing bit flips after s)."""
    common = _delta(keys, first, last)
    split = first
    step = last - first
    while True:
        step = (step + 1) >> 1
        new = split + step
        if new < last and _delta(keys, first, new) > common:
            split = new
        if step <= 1:
            break
    return split


def build_tree(positions: np.ndarray, masses: np.ndarray, leaf_size: int, bounds=None):
    """Sort + buckets + radix tree + geometry; returns a Topology without lists."""
    x = np.asarray(positions, np.float64)
    m = np.asarray(masses, np.float64)
    n = x.shape[0]
    if n < 1:
        raise ValueError("need at least one particle")
    if leaf_size < 1:
        raise ValueError("leaf_size must be >= 1")
    codes = morton_codes(x, bounds)
    perm = np.lexsort((np.arange(n), codes))
    scodes = codes[perm]
    L = (n + leaf_size - 1) // leaf_size
    I = L - 1
    nodes = 2 * L - 1
    leaf_keys = scodes[np.arange(L) * leaf_size]

    parent = np.full(nodes, -1, np.int64)
    left = np.zeros(max(I, 0), np.int64)
    right = np.zeros(max(I, 0), np.int64)
    leaf_rng = np.zeros((nodes, 2), np.int64)  # leaf index range per node
    for l in range(L):
        leaf_rng[I + l] = (l, l)

    if L > 1:
        sys.setrecursionlimit(max(10000, 4 * L))
        stack = [(0, 0, L - 1)]
        while stack:
            node, first, last = stack.pop()
            leaf_rng[node] = (first, last)
            s = _find_split(leaf_keys, first, last)
            lc = I + s if s == first else s
            rc = I + s + 1 if s + 1 == last else s + 1
            left[node] = lc
            right[node] = rc
            parent[lc] = node
            parent[rc] = node
            if lc < I:
                stack.append((lc, first, s))
            if rc < I:
                stack.append((rc, s + 1, last))

    node_range = np.stack([leaf_rng[:, 0] * leaf_size, np.minimum((leaf_rng[:, 1] + 1) * leaf_size, n) - 1], axis=1)

    depth = np.zeros(nodes, np.int64)
    order = [0]
    for node in order:
        if node < I:
            for c in (left[node], right[node]):
                depth[c] = depth[node] + 1
                order.append(int(c))

    xs = x[perm]
    ms = m[perm]
    mass = np.zeros(nodes)
    center = np.zeros((nodes, 3))
    # Sums run sequentially in tree-slot order (np.cumsum), the order the spec
    # fixes so that geometry -- and hence every MAC decision -- is reproducible.
    def seq_mean(a, b):
        return np.cumsum(xs[a : b + 1], axis=0)[-1] / float(b - a + 1)

    # leaves: direct sums over their particles
    for l in range(L):
        a, b = node_range[I + l]
        mm = ms[a : b + 1]
        tot = np.cumsum(mm)[-1]
        mass[I + l] = tot
        if tot > 0:
            center[I + l] = np.cumsum(mm[:, None] * xs[a : b + 1], axis=0)[-1] / tot
        else:
            center[I + l] = seq_mean(a, b)
    # internal: combine children, deepest first
    for node in sorted(range(I), key=lambda k: -depth[k]):
        lc, rc = left[node], right[node]
        tot = mass[lc] + mass[rc]
        mass[node] = tot
        if tot > 0:
            center[node] = (mass[lc] * center[lc] + mass[rc] * center[rc]) / tot
        else:
            a, b = node_range[node]
            center[node] = seq_mean(a, b)
    radius = np.zeros(nodes)
    for node in range(nodes):
        a, b = node_range[node]
        d = xs[a : b + 1] - center[node]
        radius[node] = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]).max())

    return Topology(
        num_particles=n, leaf_size=leaf_size, num_leaves=L, num_internal=I, num_nodes=nodes,
        perm=perm.astype(np.int64), codes=scodes, parent=parent, left=left, right=right,
        node_range=node_range, depth=depth, mass=mass, center=center, radius=radius,
        far_off=np.zeros(nodes + 1, np.int64), far_src=np.zeros(0, np.int64),
        near_off=np.zeros(L + 1, np.int64), near_src=np.zeros(0, np.int64),
    )


def mac_accepts(topo: Topology, a: int, b: int, theta: float) -> bool:
    """Sum-form MAC with a strictly positive separation."""
    d = topo.center[a] - topo.center[b]
    d2 = float(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    s = topo.radius[a] + topo.radius[b]
    return d2 > 0.0 and s * s <= theta * theta * d2


def build_lists(topo: Topology, theta: float) -> Topology:
    """Dual-tree walk from (root, root) -> far (target, source) and near leaf pairs."""
    I = topo.num_internal
    far = []
    near = []

    def is_leaf(k):
        return k >= I

    stack = [("self", 0, 0)]
    while stack:
        kind, a, b = stack.pop()
        if kind == "self":
            if is_leaf(a):
                continue
            la, ra = int(topo.left[a]), int(topo.right[a])
            stack.append(("pair", la, ra))
            stack.append(("self", ra, ra))
            stack.append(("self", la, la))
            continue
        if mac_accepts(topo, a, b, theta):
            far.append((a, b))
            far.append((b, a))
            continue
        al, bl = is_leaf(a), is_leaf(b)
        if al and bl:
            near.append((a - I, b - I))
            near.append((b - I, a - I))
            continue
        split_a = bl or ((not al) and topo.radius[a] >= topo.radius[b])
        if split_a:
            stack.append(("pair", int(topo.right[a]), b))
            stack.append(("pair", int(topo.left[a]), b))
        else:
            stack.append(("pair", a, int(topo.right[b])))
            stack.append(("pair", a, int(topo.left[b])))

    far_arr = np.array(sorted(far), dtype=np.int64).reshape(-1, 2)
    near_arr = np.array(sorted(near), dtype=np.int64).reshape(-1, 2)
    topo.far_off = np.searchsorted(far_arr[:, 0], np.arange(topo.num_nodes + 1)).astype(np.int64)
    topo.far_src = far_arr[:, 1].copy()
    topo.near_off = np.searchsorted(near_arr[:, 0], np.arange(topo.num_leaves + 1)).astype(np.int64)
    topo.near_src = near_arr[:, 1].copy()
    return topo


def build_topology(positions, masses, leaf_size: int, theta: float, bounds=None) -> Topology:
    """The full fixed topology: tree, geometry and far/near lists."""
    return build_lists(build_tree(positions, masses, leaf_size, bounds), theta)


def internal_nodes_by_depth_desc(topo: Topology):
    """Internal node ids grouped by depth, deepest level first (M2M order)."""
    I = topo.num_internal
    if I == 0:
        return []
    d = topo.depth[:I]
    return [np.nonzero(d == k)[0] for k in range(int(d.max()), -1, -1)]


def nodes_by_depth_asc(topo: Topology):
    """All non-root nodes grouped by depth, shallowest first (L2L order)."""
    d = topo.depth
    return [np.nonzero(d == k)[0] for k in range(1, int(d.max()) + 1)]
