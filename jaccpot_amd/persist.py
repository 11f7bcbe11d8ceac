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
        "topology": {k: (float(getattr(topo, k)) if k == "theta" else int(getattr(topo, k))) for k in TOPO_SCALARS},
    }
    arrays = {f"topology.{k}": np.ascontiguousarray(np.asarray(getattr(topo, k))) for k in TOPO_ARRAYS}
    for k in ("pos4", "centers4", "multipoles", "locals"):
        arrays[k] = getattr(state, k).cpu().numpy()
    if gears is not None:
        arrays["gear_of_pair"] = np.asarray(state.extra["gear_of_pair"], np.int8)
    with open(path, "wb") as f:
        np.savez(f, meta=np.array(json.dumps(meta)), **arrays)


def load_prepared_state(fmm, path):
    from .learned import build_gear_schedule

    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        if meta.get("format") != FORMAT or int(meta.get("version", -1)) != VERSION:
            raise ValueError(f"{path}: not a {FORMAT} v{VERSION} file")
        tmeta = meta["topology"]
        topo = Topology(**{k: tmeta[k] for k in TOPO_SCALARS},
                        **{k: np.array(z[f"topology.{k}"]) for k in TOPO_ARRAYS}, build_seconds=0.0)
        dtype = getattr(torch, meta["dtype"])
        if meta["input_kind"] == "numpy":
            in_dtype, in_dev = np.dtype(meta["input_dtype"]), None
        else:
            in_dtype, in_dev = getattr(torch, meta["input_dtype"]), fmm.device
        state = fmm._state_from_topology(topo, int(meta["order"]), dtype, float(meta["theta"]), meta["input_kind"],
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
            g = np.array(z["gear_of_pair"])
            state.extra["gears"] = build_gear_schedule(state, g, tuple(meta["gears"]))
            state.extra["gear_of_pair"] = g
    return state
