"""Work items for the P2P near field: long near lists cut into equal slot ranges.

A leaf's P2P sources are its own particles and those of its near leaves
(``leaf_size`` slots per leaf).  Near-list lengths are very uneven (Plummer
N=4M, leaf 8, theta 0.85: median 150 leaves, a few outer-halo leaves with
wide bounding spheres over 500,000), and one workgroup per leaf leaves a
handful of leaves running alone at the end of the launch.  Items of at most
``CHUNK_SLOTS`` slots spread every leaf's work over the GPU; a leaf that needs
several items gets per-item partial sums that ``jfm_p2p_reduce`` adds in item
order (deterministic).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

# Slots per item: measured on MI355X (Plummer N=4M, leaf 8 / 64), 512: 8.2 / 69.4 ms,
# 2048: 4.17 / 64.3 ms, 8192: 3.96 / 63.8 ms per P2P launch.
CHUNK_SLOTS = 8192


@dataclass
class P2PItems:
    num_items: int
    leaf: torch.Tensor
    s0: torch.Tensor
    s1: torch.Tensor
    part: torch.Tensor
    num_part: int
    partials: torch.Tensor
    num_split: int
    split_leaf: torch.Tensor
    split_p0: torch.Tensor
    split_p1: torch.Tensor


def build_p2p_items(near_off: np.ndarray, leaf_size: int, dtype: torch.dtype, device,
                    chunk_slots: int = CHUNK_SLOTS):
    """Items for the near CSR ``near_off`` [L+1]; None when every leaf fits one item."""
    if chunk_slots < 1:
        raise ValueError(f"chunk_slots must be positive, got {chunk_slots}")
    # item bounds are whole source leaves (the leaf-granular fetch of the P2P
    # kernel reads a leaf's slots in aligned groups)
    chunk_slots = max(chunk_slots // int(leaf_size), 1) * int(leaf_size)
    near_off = np.asarray(near_off, np.int64)
    L = near_off.size - 1
    slots = (np.diff(near_off) + 1) * int(leaf_size)
    nch = (slots + chunk_slots - 1) // chunk_slots
    if L == 0 or int(nch.max()) <= 1:
        return None
    total = int(nch.sum())
    item_leaf = np.repeat(np.arange(L, dtype=np.int64), nch)
    first = np.concatenate([[0], np.cumsum(nch)[:-1]])
    k = np.arange(total, dtype=np.int64) - np.repeat(first, nch)
    s0 = k * chunk_slots
    s1 = np.minimum(s0 + chunk_slots, np.repeat(slots, nch))
    multi = np.repeat(nch > 1, nch)
    part = np.full(total, -1, np.int64)
    part[multi] = np.arange(int(multi.sum()))
    split = np.nonzero(nch > 1)[0]
    p0 = part[first[split]]
    p1 = p0 + nch[split]
    assert int(s1.max()) < 2 ** 31 and np.all(s1 > s0)

    def dev(a):
        return torch.as_tensor(np.ascontiguousarray(a, np.int32)).to(device)

    num_part = int(multi.sum())
    return P2PItems(num_items=total, leaf=dev(item_leaf), s0=dev(s0), s1=dev(s1), part=dev(part), num_part=num_part,
                    partials=torch.empty((max(num_part, 1), int(leaf_size), 4), dtype=dtype, device=device),
                    num_split=int(split.size), split_leaf=dev(split), split_p0=dev(p0), split_p1=dev(p1))
