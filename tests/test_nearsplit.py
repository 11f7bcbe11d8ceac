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


@pytest.mark.parametrize("leaf,chunk", [(5, 64), (12, 100), (8, 20), (3, 2)])
def test_item_bounds_are_whole_leaves(leaf, chunk):
    """The chunk is rounded down to whole source leaves (at least one)."""
    from jaccpot_amd.nearsplit import build_p2p_items

    rng = np.random.default_rng(1)
    counts = rng.integers(0, 60, size=200)
    near_off = np.concatenate([[0], np.cumsum(counts)])
    it = build_p2p_items(near_off, leaf, torch.float32, "cpu", chunk_slots=chunk)
    s0, s1, leaf_of = it.s0.numpy(), it.s1.numpy(), it.leaf.numpy()
    step = max(chunk // leaf, 1) * leaf
    assert np.all(s0 % leaf == 0) and np.all(s1 % leaf == 0)
    assert np.all(s1 - s0 <= step) and np.all(s1 > s0)
    # the items of each leaf tile its whole slot sequence in order
    for l in (0, 57, 199):
        sel = np.nonzero(leaf_of == l)[0]
        assert s0[sel[0]] == 0 and s1[sel[-1]] == (counts[l] + 1) * leaf
        assert np.all(s1[sel[:-1]] == s0[sel[1:]])


def test_no_items_when_every_leaf_fits():
    from jaccpot_amd.nearsplit import build_p2p_items

    near_off = np.arange(0, 101, 10)
    assert build_p2p_items(near_off, 8, torch.float32, "cpu", chunk_slots=2048) is None


def test_prepared_state_file_holds_plain_arrays(tmp_path):
    """The on-disk state is a JSON header plus plain arrays (no pickled objects)."""
    import json
    from types import SimpleNamespace

    from jaccpot_amd.persist import FORMAT, TOPO_ARRAYS, save_prepared_state
    from jaccpot_amd.topology import build_topology
    from oracle import ics

    x, m = ics.plummer(2000, 3)
    topo = build_topology(x, m, leaf_size=16, theta=0.6)
    nodes = topo.num_nodes
    state = SimpleNamespace(topology=topo, order=4, dtype=torch.float64, G=1.0, softening=1e-3, theta=0.6,
                            compensated=True, input_kind="numpy", input_dtype=np.dtype("float64"), extra={},
                            pos4=torch.zeros((2000, 4), dtype=torch.float64),
                            centers4=torch.zeros((nodes, 4), dtype=torch.float64),
                            multipoles=torch.zeros((nodes, 28), dtype=torch.float64),
                            locals=torch.ones((nodes, 28), dtype=torch.float64))
    path = tmp_path / "state.npz"
    save_prepared_state(state, path)
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        assert meta["format"] == FORMAT and meta["order"] == 4 and meta["gears"] is None
        for k in TOPO_ARRAYS:
            assert np.array_equal(z[f"topology.{k}"], np.asarray(getattr(topo, k)))
        assert np.array_equal(z["locals"], np.ones((nodes, 28)))
