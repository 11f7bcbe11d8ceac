"""Flat (lane-per-pair) form of a far list for ``jfm_m2l_flat``.

A CSR far list (pairs grouped by target) is kept as two device arrays, the
target and the source of every pair, plus the list of targets whose pairs
cross a 64-pair block boundary (their per-block partial sums are added in
block order by the fix-up pass) and the per-block scratch rows.
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


def build_flat_list(tgt: np.ndarray, src: np.ndarray, gear: int, dtype: torch.dtype, device) -> FlatList:
    """``tgt``/``src`` [K]: target and source node of every pair, grouped by target."""
    tgt = np.ascontiguousarray(tgt, np.int32)
    src = np.ascontiguousarray(src, np.int32)
    K = int(tgt.shape[0])
    if K and np.any(np.diff(tgt) < 0):
        raise ValueError("pairs must be grouped by target in ascending order")
    span_t = span_w0 = span_w1 = np.zeros(0, np.int32)
    if K:
        starts = np.concatenate([[0], np.nonzero(np.diff(tgt))[0] + 1])
        ends = np.concatenate([starts[1:], [K]]) - 1
        w0 = starts // BLOCK
        w1 = ends // BLOCK
        sp = w1 > w0
        span_t, span_w0, span_w1 = tgt[starts[sp]], w0[sp].astype(np.int32), w1[sp].astype(np.int32)
    n_scr = int(nat.load().jfm_m2l_flat_scratch_size(int(gear), K))
    if n_scr < 0:
        raise ValueError(f"unsupported gear {gear}")

    def dev(a):
        return torch.as_tensor(a if a.size else np.zeros(1, np.int32)).to(device)

    return FlatList(gear=int(gear), num_pairs=K, pair_tgt=dev(tgt), pair_src=dev(src), num_span=int(span_t.size),
                    span_tgt=dev(span_t), span_w0=dev(span_w0), span_w1=dev(span_w1),
                    scratch=torch.empty(max(n_scr, 1), dtype=dtype, device=device))


def launch_flat(fl: FlatList, stream, code: int, p: int, multipoles: torch.Tensor, centers4: torch.Tensor,
                locals_: torch.Tensor) -> None:
    if fl.num_pairs == 0:
        return
    nat.call("jfm_m2l_flat", stream, code, p, fl.gear, fl.num_pairs, fl.pair_tgt.data_ptr(),
             fl.pair_src.data_ptr(), multipoles.data_ptr(), centers4.data_ptr(), locals_.data_ptr(),
             fl.scratch.data_ptr(), fl.num_span, fl.span_tgt.data_ptr(), fl.span_w0.data_ptr(),
             fl.span_w1.data_ptr())
