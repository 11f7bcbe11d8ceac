"""P2P work items (host logic): every leaf's slot sequence is covered exactly
once, in order, and split leaves own consecutive partial rows."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_items_cover_every_slot_once():
    from jaccpot_amd.nearsplit import build_p2p_items

    rng = np.random.default_rng(0)
    counts = rng.integers(0, 40, size=300)
    counts[[5, 77, 299]] = [500, 1000, 2000]
    near_off = np.concatenate([[0], np.cumsum(counts)])
    leaf = 8
    it = build_p2p_items(near_off, leaf, torch.float32, "cpu", chunk_slots=64)
    leaf_of = it.leaf.numpy()
    s0, s1, part = it.s0.numpy(), it.s1.numpy(), it.part.numpy()
    assert np.all(np.diff(leaf_of) >= 0)
    for l in range(counts.size):
        sel = np.nonzero(leaf_of == l)[0]
        slots = (counts[l] + 1) * leaf
        assert s0[sel[0]] == 0 and s1[sel[-1]] == slots
        assert np.all(s1[sel[:-1]] == s0[sel[1:]]) and np.all(s1[sel] - s0[sel] <= 64)
        if sel.size == 1:
            assert part[sel[0]] == -1
        else:
            assert np.array_equal(part[sel], np.arange(part[sel[0]], part[sel[0]] + sel.size))
    split = it.split_leaf.numpy()
    p0, p1 = it.split_p0.numpy(), it.split_p1.numpy()
    assert np.array_equal(split, np.nonzero((counts + 1) * leaf > 64)[0])
    assert p0[0] == 0 and np.all(p1[:-1] == p0[1:]) and p1[-1] == it.num_part
    assert it.partials.shape == (it.num_part, leaf, 4)


def test_no_items_when_every_leaf_fits():
    from jaccpot_amd.nearsplit import build_p2p_items

    near_off = np.arange(0, 101, 10)
    assert build_p2p_items(near_off, 8, torch.float32, "cpu", chunk_slots=2048) is None
