"""Flat (lane-per-pair) form of a far list for ``jfm_m2l_flat``.

A CSR far list (pairs grouped by target) is kept as two device arrays, the
target and the source of every pair, plus the list of targets whose pairs
cross a 64-pair block boundary (their per-block partial sums are added in
block order by the fix-up pass; each such target is handled on its own, so
the order of that list does not change any result) and the per-block
scratch rows.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat

BLOCK = 64  # pairs per wave in jfm_m2l_flat

# A far list runs lane-per-pair when its targets average fewer pairs than
# this: the wave-per-target kernel then idles many lanes, while for denser
# lists it keeps its lanes busy and needs no cross-wave fix-up (measured on
# MI355X: gear-2/4 lists with ~10 pairs per target run 2x faster flat, a
# gear-8 list with ~36 pairs per target costs 15% more per pair than the
# same gear with ~57, the gear-6 list with ~80 pairs per target runs 4%
# slower flat plus the fix-up pass).
FLAT_MAX_MEAN_PAIRS = 56


def prefers_flat(num_pairs: int, num_targets: int) -> bool:
    return num_pairs > 0 and num_pairs < FLAT_MAX_MEAN_PAIRS * max(num_targets, 1)


@dataclass
class FlatList:
    gear: int
    num_pairs: int
    pair_tgt: torch.Tensor
    pair_src: torch.Tensor
    num_span: int
    span_tgt: torch.Tensor
    span_w0: torch.Tensor
    span_w1: torch.Tensor
    scratch: torch.Tensor


def _as_device_i32(a, device) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.int32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, np.int32)).to(device)


def build_flat_list(tgt, src, gear: int, dtype: torch.dtype, device) -> FlatList:
    """``tgt``/``src`` [K] (numpy arrays or device tensors): target and source node
    of every pair, grouped by target in ascending order.  The spanning targets
    are found on the device (``jfm_flat_spans``)."""
    pair_tgt = _as_device_i32(tgt, device)
    pair_src = _as_device_i32(src, device)
    K = int(pair_tgt.numel())
    if K != int(pair_src.numel()):
        raise ValueError("tgt and src must have the same length")
    if K > 1 and bool((pair_tgt[1:] < pair_tgt[:-1]).any()):
        raise ValueError("pairs must be grouped by target in ascending order")
    n_scr = int(nat.load().jfm_m2l_flat_scratch_size(int(gear), K))
    if n_scr < 0:
        raise ValueError(f"unsupported gear {gear}")
    cap = max((K - 1) // BLOCK, 0) + 1
    span_tgt = torch.empty(cap, dtype=torch.int32, device=device)
    span_w0 = torch.empty(cap, dtype=torch.int32, device=device)
    span_w1 = torch.empty(cap, dtype=torch.int32, device=device)
    count = torch.zeros(1, dtype=torch.int64, device=device)
    if K:
        nat.call("jfm_flat_spans", nat.current_stream(), K, pair_tgt.data_ptr(), span_tgt.data_ptr(),
                 span_w0.data_ptr(), span_w1.data_ptr(), count.data_ptr())
    num_span = int(count.item())
    one = torch.zeros(1, dtype=torch.int32, device=device)
    return FlatList(gear=int(gear), num_pairs=K, pair_tgt=pair_tgt if K else one, pair_src=pair_src if K else one,
                    num_span=num_span, span_tgt=span_tgt, span_w0=span_w0, span_w1=span_w1,
                    scratch=torch.empty(max(n_scr, 1), dtype=dtype, device=device))


def pair_targets(state) -> torch.Tensor:
    """Target node of every far pair of ``state`` (device int32, CSR order)."""
    K = state.topology.num_far_pairs
    row_of = torch.empty(max(K, 1), dtype=torch.int32, device=state.device)
    if K:
        nat.call("jfm_pair_rows", nat.current_stream(), state.topology.num_nodes, state.far_off.data_ptr(), K,
                 row_of.data_ptr())
    return row_of[:K]


def launch_flat(fl: FlatList, stream, code: int, p: int, multipoles: torch.Tensor, centers4: torch.Tensor,
                locals_: torch.Tensor) -> None:
    if fl.num_pairs == 0:
        return
    nat.call("jfm_m2l_flat", stream, code, p, fl.gear, fl.num_pairs, fl.pair_tgt.data_ptr(),
             fl.pair_src.data_ptr(), multipoles.data_ptr(), centers4.data_ptr(), locals_.data_ptr(),
             fl.scratch.data_ptr(), fl.num_span, fl.span_tgt.data_ptr(), fl.span_w0.data_ptr(),
             fl.span_w1.data_ptr())
