"""Topology built on the GPU: the host builder's spec, bit-identical output.

Same role as ``topology.build_topology`` (yggdrax's tree and interaction-list
builders for the reference, jaccpot/runtime/fmm_state.py:630-800,
runtime/_interaction_cache.py:523,564), run as libjfm device kernels
(``csrc/treebuild.hip``): bounds, Morton codes, stable radix sort, radix tree,
levels, centres of mass and radii, breadth-first dual-tree walk, CSR lists.
Returns a ``Topology`` (host arrays, so every host consumer works unchanged;
the three largest are copied from the device on first read) with the device
arrays in ``extra["device"]`` so that ``prepare_state`` does not upload them
again.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _native as nat
from .topology import Topology


class _Lazy:
    """A device array whose host copy is made on first access."""

    __slots__ = ("t",)

    def __init__(self, t: torch.Tensor):
        self.t = t


class DeviceTopology(Topology):
    """``Topology`` whose largest host arrays (perm, far_src, near_src) are copied
    from the device only when first read; the device build never needs them."""

    def __getattribute__(self, name):
        v = object.__getattribute__(self, name)
        if type(v) is _Lazy:
            v = v.t.cpu().numpy()
            object.__setattr__(self, name, v)
        return v

    @property
    def num_far_pairs(self) -> int:
        return int(self.extra["num_far_pairs"])

    @property
    def num_near_entries(self) -> int:
        return int(self.extra["num_near_entries"])


def _sort_u64(keys: torch.Tensor, end_bit: int, vals: torch.Tensor | None = None):
    n = int(keys.numel())
    out_k = torch.empty_like(keys)
    out_v = torch.empty_like(vals) if vals is not None else None
    if n == 0:
        return out_k, out_v
    nb = ctypes.c_int64(0)
    nat.call("jfm_sort_scratch_bytes", n, 1 if vals is not None else 0, ctypes.byref(nb))
    scratch = torch.empty(max(int(nb.value), 1), dtype=torch.uint8, device=keys.device)
    nat.call("jfm_sort_u64", nat.current_stream(), n, int(end_bit), keys.data_ptr(), out_k.data_ptr(),
             nat.ptr(vals), nat.ptr(out_v), scratch.data_ptr(), int(nb.value))
    return out_k, out_v


def _csr(sorted_keys: torch.Tensor, rows: int):
    dev = sorted_keys.device
    off = torch.empty(rows + 1, dtype=torch.int32, device=dev)
    col = torch.empty(max(int(sorted_keys.numel()), 1), dtype=torch.int32, device=dev)
    nat.call("jfm_keys_to_csr", nat.current_stream(), nat.ptr(sorted_keys) if sorted_keys.numel() else None,
             int(sorted_keys.numel()), int(rows), off.data_ptr(), col.data_ptr())
    return off, col[: int(sorted_keys.numel())]


def _key_bits(max_key: int) -> int:
    return max(1, int(nat.call("jfm_key_bits", ctypes.c_uint64(int(max_key)))))


class _Grow:
    """A device buffer that keeps its first ``used`` entries when it grows."""

    def __init__(self, cap: int, dtype, device, width: int = 1):
        self.width = width
        self.buf = torch.empty((max(cap, 1), width) if width > 1 else (max(cap, 1),), dtype=dtype, device=device)

    @property
    def cap(self) -> int:
        return int(self.buf.shape[0])

    def grow(self, need: int, used: int) -> None:
        new_cap = max(need, int(self.cap * 1.5) + 1)
        nb = torch.empty((new_cap,) + tuple(self.buf.shape[1:]), dtype=self.buf.dtype, device=self.buf.device)
        if used:
            nb[:used].copy_(self.buf[:used])
        self.buf = nb


def build_topology_device(x: torch.Tensor, m: torch.Tensor, *, leaf_size: int, theta: float,
                          bounds=None) -> Topology:
    """Build the topology for input-order float64 device tensors x [N,3], m [N]."""
    if x.dtype != torch.float64 or m.dtype != torch.float64 or x.device.type != "cuda":
        raise ValueError("build_topology_device needs float64 CUDA tensors")
    if x.ndim != 2 or x.shape[1] != 3 or m.shape != (x.shape[0],):
        raise ValueError("positions must be [N,3] and masses [N]")
    n = int(x.shape[0])
    if n < 1 or n > (1 << 30):
        raise ValueError("need 1 <= N <= 2^30 particles")
    if leaf_size < 1 or leaf_size > 256:
        raise ValueError("leaf_size must be in [1, 256]")
    if not (theta > 0.0) or not np.isfinite(theta):
        raise ValueError("theta must be > 0")
    x = x.contiguous()
    m = m.contiguous()
    dev = x.device
    s = nat.current_stream()
    t0 = time.perf_counter()
    L = (n + leaf_size - 1) // leaf_size
    I = L - 1
    nodes = 2 * L - 1

    # ---- bounds, Morton codes, sort ------------------------------------------------------
    flags = torch.zeros(1, dtype=torch.int32, device=dev)
    bscratch = torch.empty(int(nat.call("jfm_tree_bounds_scratch_doubles")), dtype=torch.float64, device=dev)
    if bounds is not None:
        lo, hi = bounds
        b6 = torch.as_tensor(np.concatenate([np.asarray(lo, np.float64), np.asarray(hi, np.float64)]), device=dev)
    else:
        b6 = torch.empty(6, dtype=torch.float64, device=dev)
    nat.call("jfm_tree_bounds", s, x.data_ptr(), n, 1 if bounds is not None else 0, bscratch.data_ptr(),
             b6.data_ptr(), flags.data_ptr())
    codes = torch.empty(n, dtype=torch.int64, device=dev)
    index = torch.empty(n, dtype=torch.int32, device=dev)
    nat.call("jfm_tree_morton", s, x.data_ptr(), n, b6.data_ptr(), codes.data_ptr(), index.data_ptr())
    sorted_codes, perm = _sort_u64(codes, 63, index)
    del codes, index
    bounds_used = b6

    # ---- radix tree, levels ---------------------------------------------------------------
    parent = torch.empty(nodes, dtype=torch.int32, device=dev)
    left = torch.empty(max(I, 1), dtype=torch.int32, device=dev)
    right = torch.empty(max(I, 1), dtype=torch.int32, device=dev)
    node_range = torch.empty((nodes, 2), dtype=torch.int32, device=dev)
    nat.call("jfm_tree_radix", s, n, int(leaf_size), sorted_codes.data_ptr(), parent.data_ptr(), left.data_ptr(),
             right.data_ptr(), node_range.data_ptr())
    depth = torch.empty(nodes, dtype=torch.int32, device=dev)
    maxd = torch.zeros(1, dtype=torch.int32, device=dev)
    m2m_keys = torch.empty(max(I, 1), dtype=torch.int64, device=dev)
    l2l_keys = torch.empty(max(nodes - 1, 1), dtype=torch.int64, device=dev)
    nat.call("jfm_tree_levels", s, nodes, I, parent.data_ptr(), depth.data_ptr(), maxd.data_ptr(), flags.data_ptr(),
             m2m_keys.data_ptr(), l2l_keys.data_ptr())
    hdr = torch.cat([flags, maxd]).cpu().numpy()
    if hdr[0] & 1:
        raise ValueError("positions must be finite")
    if hdr[0] & 2:
        raise RuntimeError("device topology build: parent chain exceeds the maximum tree height")
    D = int(hdr[1])
    if I > 0:
        sk, _ = _sort_u64(m2m_keys[:I], 32 + _key_bits(D))
        m2m_off_d, m2m_nodes = _csr(sk, D + 1)
        sk, _ = _sort_u64(l2l_keys[: nodes - 1], 32 + _key_bits(D))
        l2l_off_d, l2l_nodes = _csr(sk, D)
        m2m_off = m2m_off_d.cpu().numpy()
        l2l_off = l2l_off_d.cpu().numpy()
    else:
        m2m_nodes = torch.empty(0, dtype=torch.int32, device=dev)
        l2l_nodes = torch.empty(0, dtype=torch.int32, device=dev)
        m2m_off = np.zeros(1, np.int32)
        l2l_off = np.zeros(1, np.int32)
    del m2m_keys, l2l_keys

    # ---- geometry ------------------------------------------------------------------------
    mass = torch.empty(nodes, dtype=torch.float64, device=dev)
    center = torch.empty((nodes, 3), dtype=torch.float64, device=dev)
    radius = torch.empty(nodes, dtype=torch.float64, device=dev)
    nat.call("jfm_tree_leaf_geometry", s, L, I, x.data_ptr(), m.data_ptr(), perm.data_ptr(), node_range.data_ptr(),
             mass.data_ptr(), center.data_ptr())
    for lev in range(len(m2m_off) - 1):
        a, b = int(m2m_off[lev]), int(m2m_off[lev + 1])
        if b > a:
            nat.call("jfm_tree_level_geometry", s, b - a, m2m_nodes[a:].data_ptr(), x.data_ptr(), m.data_ptr(),
                     perm.data_ptr(), left.data_ptr(), right.data_ptr(), node_range.data_ptr(), mass.data_ptr(),
                     center.data_ptr())
    xs = torch.empty((n, 3), dtype=torch.float64, device=dev)
    nat.call("jfm_tree_radius", s, nodes, n, x.data_ptr(), perm.data_ptr(), node_range.data_ptr(), center.data_ptr(),
             xs.data_ptr(), radius.data_ptr())
    del xs

    # ---- dual-tree walk, breadth first: classify, scan, emit ------------------------------
    items = torch.zeros((1, 2), dtype=torch.int32, device=dev)  # the root self item (0, 0)
    far = _Grow(max(1024, 128 * L), torch.int64, dev)
    near = _Grow(max(1024, 320 * L), torch.int64, dev)
    n_in, n_far, n_near = 1, 0, 0
    scan_bytes = ctypes.c_int64(0)
    scratch = torch.empty(1, dtype=torch.uint8, device=dev)
    while n_in > 0:
        codes = torch.empty(n_in, dtype=torch.uint8, device=dev)
        counts = torch.empty((3, n_in), dtype=torch.int32, device=dev)
        offsets = torch.empty((3, n_in), dtype=torch.int32, device=dev)
        nat.call("jfm_tree_walk_classify", s, I, items.data_ptr(), n_in, center.data_ptr(), radius.data_ptr(),
                 float(theta), codes.data_ptr(), counts.data_ptr())
        nat.call("jfm_scan_scratch_bytes", n_in, ctypes.byref(scan_bytes))
        if scratch.numel() < scan_bytes.value:
            scratch = torch.empty(int(scan_bytes.value), dtype=torch.uint8, device=dev)
        for r in range(3):
            nat.call("jfm_scan_i32", s, n_in, counts[r].data_ptr(), offsets[r].data_ptr(), scratch.data_ptr(),
                     int(scratch.numel()))
        totals = (offsets[:, -1] + counts[:, -1]).cpu().numpy()
        t_items, t_far, t_near = (int(v) for v in totals)
        if n_far + t_far > far.cap:
            far.grow(n_far + t_far, n_far)
        if n_near + t_near > near.cap:
            near.grow(n_near + t_near, n_near)
        nxt = torch.empty((max(t_items, 1), 2), dtype=torch.int32, device=dev)
        nat.call("jfm_tree_walk_emit", s, I, items.data_ptr(), n_in, left.data_ptr(), right.data_ptr(),
                 codes.data_ptr(), offsets.data_ptr(), n_far, n_near, nxt.data_ptr(), far.buf.data_ptr(),
                 near.buf.data_ptr())
        items, n_in = nxt, t_items
        n_far += t_far
        n_near += t_near
    del items, codes, counts, offsets, scratch
    fk, _ = _sort_u64(far.buf[:n_far], 32 + _key_bits(max(nodes - 1, 1)))
    del far
    far_off, far_src = _csr(fk, nodes)
    del fk
    nk, _ = _sort_u64(near.buf[:n_near], 32 + _key_bits(max(L - 1, 1)))
    del near
    near_off, near_src = _csr(nk, L)
    del nk
    torch.cuda.synchronize(dev)
    device = dict(perm=perm, parent=parent, left=left, right=right, node_range=node_range, depth=depth,
                  center=center, radius=radius, mass=mass, far_off=far_off, far_src=far_src, near_off=near_off,
                  near_src=near_src, m2m_nodes=m2m_nodes, l2l_nodes=l2l_nodes, sorted_codes=sorted_codes,
                  bounds=bounds_used)
    t1 = time.perf_counter()

    def h(t):
        return t.cpu().numpy()

    far_off_h = h(far_off)
    topo = DeviceTopology(
        num_particles=n, leaf_size=int(leaf_size), num_leaves=L, num_internal=I, num_nodes=nodes, max_depth=D,
        theta=float(theta), perm=_Lazy(perm), parent=h(parent), left=h(left[:I]), right=h(right[:I]),
        node_range=h(node_range), depth=h(depth), center=h(center), radius=h(radius), mass=h(mass),
        far_off=far_off_h, far_src=_Lazy(far_src), near_off=h(near_off), near_src=_Lazy(near_src),
        m2m_nodes=h(m2m_nodes), m2m_off=m2m_off.astype(np.int32), l2l_nodes=h(l2l_nodes),
        l2l_off=l2l_off.astype(np.int32),
        m2l_targets=np.nonzero(np.diff(far_off_h) > 0)[0].astype(np.int32),
        build_seconds=0.0,
    )
    topo.extra["num_far_pairs"] = n_far
    topo.extra["num_near_entries"] = n_near
    topo.build_seconds = time.perf_counter() - t0
    topo.extra["device"] = device
    topo.extra["device_build_seconds"] = t1 - t0
    return topo


def sorted_morton(x: torch.Tensor, bounds=None):
    """(sorted Morton codes, perm) of float64 device positions x [N,3]: the first
    steps of ``build_topology_device`` (the topology reuse check)."""
    n = int(x.shape[0])
    dev = x.device
    s = nat.current_stream()
    x = x.contiguous()
    flags = torch.zeros(1, dtype=torch.int32, device=dev)
    bscratch = torch.empty(int(nat.call("jfm_tree_bounds_scratch_doubles")), dtype=torch.float64, device=dev)
    if bounds is not None:
        lo, hi = bounds
        b6 = torch.as_tensor(np.concatenate([np.asarray(lo, np.float64), np.asarray(hi, np.float64)]), device=dev)
    else:
        b6 = torch.empty(6, dtype=torch.float64, device=dev)
    nat.call("jfm_tree_bounds", s, x.data_ptr(), n, 1 if bounds is not None else 0, bscratch.data_ptr(),
             b6.data_ptr(), flags.data_ptr())
    codes = torch.empty(n, dtype=torch.int64, device=dev)
    index = torch.empty(n, dtype=torch.int32, device=dev)
    nat.call("jfm_tree_morton", s, x.data_ptr(), n, b6.data_ptr(), codes.data_ptr(), index.data_ptr())
    sc, perm = _sort_u64(codes, 63, index)
    if int(flags.item()) & 1:
        raise ValueError("positions must be finite")
    return sc, perm
