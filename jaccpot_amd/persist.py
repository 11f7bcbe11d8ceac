"""On-disk form of a prepared state (the reference's ``FMMPreparedState``,
runtime/fmm_state.py:802-1073, kept across processes).

One ``.npz`` file: a JSON header (``meta``) and plain arrays -- the host
topology, the tree-order particles, node centres, multipoles and locals as
computed, and the per-pair gears of a learned-order state.  Loading uses
``numpy.load(allow_pickle=False)``; the schedules derived from the topology
(level lists, lane-per-pair lists, P2P work items, float64 top-level shadows,
per-gear lists) are rebuilt on load, so the format holds only primary data.
"""

from __future__ import annotations

import json

import numpy as np
import torch

from .topology import Topology

FORMAT = "jaccpot_amd.prepared_state"
VERSION = 1
TOPO_ARRAYS = ("perm", "parent", "left", "right", "node_range", "depth", "center", "radius", "mass", "far_off",
               "far_src", "near_off", "near_src", "m2m_nodes", "m2m_off", "l2l_nodes", "l2l_off", "m2l_targets")
TOPO_SCALARS = ("num_particles", "leaf_size", "num_leaves", "num_internal", "num_nodes", "max_depth", "theta")


def _dtype_name(d) -> str:
    return str(d).replace("torch.", "")


def save_prepared_state(state, path) -> None:
    topo = state.topology
    gears = state.extra.get("gears")
    meta = {
        "format": FORMAT, "version": VERSION, "order": state.order, "dtype": _dtype_name(state.dtype),
        "G": state.G, "softening": state.softening, "theta": state.theta, "compensated": state.compensated,
        "input_kind": state.input_kind, "input_dtype": _dtype_name(state.input_dtype),
        "mixed": "mixed" in state.extra, "split_near": state.extra.get("p2p_items") is not None,
        "gears": list(gears.gears) if gears is not None else None,
        "order_judge": state.extra.get("order_judge"),
        "topology": {k: (float(getattr(topo, k)) if k == "theta" else int(getattr(topo, k))) for k in TOPO_SCALARS},
    }
    arrays = {f"topology.{k}": np.ascontiguousarray(np.asarray(getattr(topo, k))) for k in TOPO_ARRAYS}
    for k in ("pos4", "centers4", "multipoles", "locals"):
        arrays[k] = getattr(state, k).cpu().numpy()
    if gears is not None:
        arrays["gear_of_pair"] = np.asarray(state.extra["gear_of_pair"], np.int8)
    with open(path, "wb") as f:
        np.savez(f, meta=np.array(json.dumps(meta)), **arrays)


_WORK_DTYPES = {"float32": torch.float32, "float64": torch.float64}
_INPUT_DTYPES = ("float16", "bfloat16", "float32", "float64", "int32", "int64")


def _check_csr(path, name, off, col, rows, lo, hi):
    if off.shape != (rows + 1,) or off[0] != 0 or np.any(np.diff(off) < 0) or off[-1] != col.shape[0]:
        raise ValueError(f"{path}: {name} offsets are not a CSR of {rows} rows over {col.shape[0]} entries")
    _check_range(path, name, col, lo, hi)


def _check_range(path, name, a, lo, hi):
    if a.size and (int(a.min()) < lo or int(a.max()) >= hi):
        raise ValueError(f"{path}: {name} holds indices outside [{lo}, {hi})")


def validate_topology(topo: Topology, path="topology") -> None:
    """Every index array of ``topo`` in range and every CSR well formed, so the
    kernels that use them as device indices stay inside their buffers."""
    n, leaf = int(topo.num_particles), int(topo.leaf_size)
    if n < 1 or not (1 <= leaf <= 256):
        raise ValueError(f"{path}: bad particle count {n} or leaf size {leaf}")
    L = -(-n // leaf)
    I, nodes = L - 1, 2 * L - 1
    if (topo.num_leaves, topo.num_internal, topo.num_nodes) != (L, I, nodes):
        raise ValueError(f"{path}: node counts {topo.num_leaves}/{topo.num_internal}/{topo.num_nodes} do not "
                         f"match N={n}, leaf size {leaf}")
    for k in TOPO_ARRAYS:
        a = np.asarray(getattr(topo, k))
        want = np.float64 if k in ("center", "radius", "mass") else np.int32
        if a.dtype != want:
            raise ValueError(f"{path}: {k} has dtype {a.dtype}, expected {np.dtype(want)}")
    perm = np.asarray(topo.perm)
    if perm.shape != (n,) or not np.array_equal(np.sort(perm), np.arange(n, dtype=np.int32)):
        raise ValueError(f"{path}: perm is not a permutation of the {n} particles")
    parent = np.asarray(topo.parent)
    if parent.shape != (nodes,) or parent[0] != -1:
        raise ValueError(f"{path}: parent must have {nodes} entries with -1 at the root")
    _check_range(path, "parent", parent[1:], 0, max(I, 1))
    for k in ("left", "right"):
        a = np.asarray(getattr(topo, k))
        if a.shape != (I,):
            raise ValueError(f"{path}: {k} must have {I} entries")
        _check_range(path, k, a, 1, nodes)
    rng = np.asarray(topo.node_range)
    if rng.shape != (nodes, 2) or np.any(rng[:, 0] > rng[:, 1]):
        raise ValueError(f"{path}: node_range must be [{nodes}, 2] with first <= last")
    _check_range(path, "node_range", rng, 0, n)
    depth = np.asarray(topo.depth)
    if depth.shape != (nodes,) or not (0 <= int(topo.max_depth) < nodes):
        raise ValueError(f"{path}: depth must have {nodes} entries")
    _check_range(path, "depth", depth, 0, int(topo.max_depth) + 1)
    for k, shape in (("center", (nodes, 3)), ("radius", (nodes,)), ("mass", (nodes,))):
        a = np.asarray(getattr(topo, k))
        if a.shape != shape or not np.all(np.isfinite(a)):
            raise ValueError(f"{path}: {k} must be finite with shape {shape}")
    _check_csr(path, "far list", np.asarray(topo.far_off), np.asarray(topo.far_src), nodes, 0, nodes)
    _check_csr(path, "near list", np.asarray(topo.near_off), np.asarray(topo.near_src), L, 0, L)
    for name, off, col, lo in (("M2M schedule", topo.m2m_off, topo.m2m_nodes, 0),
                               ("L2L schedule", topo.l2l_off, topo.l2l_nodes, 1)):
        off = np.asarray(off)
        if not (1 <= off.shape[0] <= int(topo.max_depth) + 2):
            raise ValueError(f"{path}: {name} has {off.shape[0] - 1} levels for depth {topo.max_depth}")
        _check_csr(path, name, off, np.asarray(col), off.shape[0] - 1, lo, max(I, 1) if lo == 0 else nodes)
    if np.asarray(topo.m2m_nodes).shape != (I,) or np.asarray(topo.l2l_nodes).shape != (nodes - 1,):
        raise ValueError(f"{path}: level schedules must list every internal / non-root node once")
    _check_range(path, "m2l_targets", np.asarray(topo.m2l_targets), 0, nodes)


def load_prepared_state(fmm, path):
    from . import _native as nat
    from .learned import build_gear_schedule

    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        if meta.get("format") != FORMAT or int(meta.get("version", -1)) != VERSION:
            raise ValueError(f"{path}: not a {FORMAT} v{VERSION} file")
        tmeta = meta["topology"]
        topo = Topology(**{k: (float(tmeta[k]) if k == "theta" else int(tmeta[k])) for k in TOPO_SCALARS},
                        **{k: np.array(z[f"topology.{k}"]) for k in TOPO_ARRAYS}, build_seconds=0.0)
        validate_topology(topo, path)
        if meta["dtype"] not in _WORK_DTYPES:
            raise ValueError(f"{path}: working dtype {meta['dtype']!r} is not float32 or float64")
        dtype = _WORK_DTYPES[meta["dtype"]]
        order = int(meta["order"])
        if not (1 <= order <= nat.MAX_ORDER):
            raise ValueError(f"{path}: order {order} outside [1, {nat.MAX_ORDER}]")
        if meta["input_kind"] not in ("numpy", "torch") or meta["input_dtype"] not in _INPUT_DTYPES:
            raise ValueError(f"{path}: unsupported input kind/dtype {meta['input_kind']}/{meta['input_dtype']}")
        if meta["input_kind"] == "numpy":
            in_dtype, in_dev = np.dtype(meta["input_dtype"]), None
        else:
            in_dtype, in_dev = getattr(torch, meta["input_dtype"]), fmm.device
        state = fmm._state_from_topology(topo, order, dtype, float(meta["theta"]), meta["input_kind"],
                                         in_dtype, in_dev, mixed=bool(meta["mixed"]),
                                         split_near=bool(meta["split_near"]), G=float(meta["G"]),
                                         softening=float(meta["softening"]), compensated=bool(meta["compensated"]))
        for k in ("pos4", "centers4", "multipoles", "locals"):
            t = getattr(state, k)
            src = torch.as_tensor(np.array(z[k]))
            if tuple(src.shape) != tuple(t.shape) or src.dtype != t.dtype:
                raise ValueError(f"{path}: {k} has shape {tuple(src.shape)} {src.dtype}, expected {tuple(t.shape)}")
            t.copy_(src.to(fmm.device))
        if meta["gears"] is not None:
            gears = tuple(int(v) for v in meta["gears"])
            g = np.array(z["gear_of_pair"])
            if (len(gears) != 3 or max(gears) != order or g.shape != (topo.num_far_pairs,)
                    or not np.all(np.isin(g, gears))):
                raise ValueError(f"{path}: per-pair gears do not match the far list or the gears {gears}")
            state.extra["gears"] = build_gear_schedule(state, g, gears)
            state.extra["gear_of_pair"] = g
        judge = meta.get("order_judge")
        if judge is not None:
            if not isinstance(judge, dict):
                raise ValueError(f"{path}: order_judge must be a mapping")
            state.extra["order_judge"] = judge
    return state
