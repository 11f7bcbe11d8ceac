"""Momentum-conserving (mutual) FMM lane with block-step level weights.

Counterpart of the reference's ``jaccpot.mutual`` (mutual/__init__.py,
topology.py, force.py) and of its block-step adapter
(nornax_adapter.py:79-1212, ``BlockStepFMM``), on libjfm kernels:

* the topology is this package's symmetric dual-tree walk (symmetric MAC
  ``(r_a + r_b)^2 <= theta^2 d^2`` on centre-of-mass radii, the larger node
  split, every unordered pair visited once and emitted in both directions);
  its canonical halves (a < b) are the reference's mutual lists
  (mutual/topology.py:300-395);
* near field: ``jfm_mutual_p2p`` -- each pair force ``G m_i m_j w d / r^3``
  is evaluated from both ends with bit-negated results (no scatter, no
  atomics), weighted by ``w[max(rung_i, rung_j)]``; long near lists are cut
  into work items (``jfm_mutual_p2p_items``, the P2P's ``nearsplit`` items)
  whose partial sums are added in item order;
* far field: P2M, M2M, then the flat M2L over the symmetric far list with
  each pair scaled by ``w[max(rung_A, rung_B)]`` (cell rung = maximum over
  the cell's particles; ``jfm_m2l_flat_weighted``), L2L and L2P; the
  order-p far force is the gradient of a pair energy symmetric in the two
  cells, so the far momentum cancels algebraically;
* forces are accumulated and divided by the mass once (``jfm_mutual_combine``).

Everything runs at a frozen topology on live positions and masses, as the
reference's mutual state does; torch inputs that require grad are
differentiated by the fixed-order adjoint kernels with the level weights on
every pair (positions and masses; topology, rungs and weights held fixed).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, NamedTuple, Optional, Tuple

import numpy as np
import torch

from . import _native as nat

__all__ = [
    "MutualTopology",
    "MutualCapacities",
    "MutualCapacityOverflow",
    "MutualFMMState",
    "MutualForceResult",
    "build_mutual_topology",
    "build_mutual_topology_from_state",
    "build_mutual_state",
    "mutual_near_field_forces",
    "mutual_far_field_forces",
    "mutual_accelerations",
    "mutual_level_accelerations",
    "mutual_weighted_accelerations",
    "boundary_level_weights",
    "boundary_weight_table",
    "level_weights_from_floor",
    "active_level_floor",
    "is_sync_boundary",
    "n_sub",
    "BlockStepFMM",
]

MAX_LEVELS = 64  # weights per call (jfm_mutual_p2p's LDS table)


# --------------------------------------------------------------------- topology
@dataclass(frozen=True)
class MutualTopology:
    """Frozen canonical pair lists of one tree (mutual/topology.py:53-185).

    Node ids ``[0, num_internal)`` are internal, ``[num_internal, num_nodes)``
    leaves; ``far_a < far_b`` are node ids, ``near_a < near_b`` leaf node ids.
    ``node_particle_ranges`` are inclusive spans of the Morton-sorted order;
    ``inverse_permutation`` maps sorted to input order.  ``prepared`` is the
    prepared FMM state whose device buffers the mutual state runs on.
    """

    num_particles: int
    num_nodes: int
    num_internal: int
    max_leaf_size: int
    order: int
    theta: float
    left_child: np.ndarray
    right_child: np.ndarray
    level_nodes: Tuple[np.ndarray, ...]
    parent_of_level_nodes: Tuple[np.ndarray, ...]
    leaf_nodes: np.ndarray
    leaf_particles: np.ndarray
    leaf_particle_valid: np.ndarray
    far_a: np.ndarray
    far_b: np.ndarray
    near_a: np.ndarray
    near_b: np.ndarray
    node_particle_ranges: np.ndarray
    inverse_permutation: np.ndarray
    forward_permutation: np.ndarray
    prepared: Any = field(default=None, repr=False, compare=False)

    @property
    def num_leaves(self) -> int:
        return int(self.leaf_nodes.shape[0])

    @property
    def num_far_pairs(self) -> int:
        return int(self.far_a.shape[0])

    @property
    def num_near_pairs(self) -> int:
        return int(self.near_a.shape[0])

    def summary(self) -> dict:
        return {"num_particles": self.num_particles, "num_nodes": self.num_nodes, "num_leaves": self.num_leaves,
                "max_leaf_size": self.max_leaf_size, "num_far_pairs": self.num_far_pairs,
                "num_near_pairs": self.num_near_pairs, "num_levels": len(self.level_nodes), "theta": self.theta,
                "order": self.order}


def build_mutual_topology_from_state(state) -> MutualTopology:
    """Canonical lists of a prepared state's symmetric topology
    (counterpart of build_mutual_topology_from_tree, mutual/topology.py:459-565)."""
    topo = state.topology
    I, L, nodes = topo.num_internal, topo.num_leaves, topo.num_nodes
    tgt = np.repeat(np.arange(nodes, dtype=np.int64), np.diff(topo.far_off))
    src = topo.far_src.astype(np.int64)
    keep = tgt < src
    far_a, far_b = tgt[keep], src[keep]
    if 2 * far_a.shape[0] != src.shape[0]:
        raise ValueError("the far list is not symmetric")
    lt = np.repeat(np.arange(L, dtype=np.int64), np.diff(topo.near_off))
    ls = topo.near_src.astype(np.int64)
    keep = lt < ls
    near_a, near_b = lt[keep] + I, ls[keep] + I
    depth = topo.depth.astype(np.int64)
    level_nodes, level_parents = [], []
    for d in range(1, int(depth.max()) + 1 if nodes > 1 else 1):
        u = np.flatnonzero(depth == d)
        if u.size:
            level_nodes.append(u)
            level_parents.append(topo.parent[u].astype(np.int64))
    leaf_nodes = np.arange(I, nodes, dtype=np.int64)
    rng = topo.node_range.astype(np.int64)
    counts = rng[leaf_nodes, 1] - rng[leaf_nodes, 0] + 1
    S = int(max(counts.max(), 1))
    off = np.arange(S)[None, :]
    valid = off < counts[:, None]
    particles = np.where(valid, rng[leaf_nodes, 0][:, None] + off, 0)
    perm = topo.perm.astype(np.int64)
    return MutualTopology(
        num_particles=topo.num_particles, num_nodes=nodes, num_internal=I, max_leaf_size=S, order=state.order,
        theta=float(topo.theta), left_child=topo.left.astype(np.int64), right_child=topo.right.astype(np.int64),
        level_nodes=tuple(level_nodes), parent_of_level_nodes=tuple(level_parents), leaf_nodes=leaf_nodes,
        leaf_particles=particles, leaf_particle_valid=valid, far_a=far_a, far_b=far_b, near_a=near_a,
        near_b=near_b, node_particle_ranges=rng, inverse_permutation=perm,
        forward_permutation=np.argsort(perm, kind="stable").astype(np.int64), prepared=state)


def build_mutual_topology(positions, masses, *, theta: float = 0.6, order: int = 4, leaf_size: int = 16,
                          solver: Optional[Any] = None) -> Tuple[MutualTopology, Any]:
    """Tree + symmetric lists with this package's builder; returns
    ``(topology, prepared_state)`` like mutual/topology.py:567-612.  The working
    precision follows the input dtype (float64 input: float64) unless ``solver``
    fixes it."""
    from .solver import FastMultipoleMethod

    if solver is None:
        solver = FastMultipoleMethod(preset="balanced", basis="real", theta=float(theta))
    if getattr(solver, "adaptive_order", False) or getattr(solver, "mixed_precision", False):
        raise NotImplementedError("the mutual lane runs at a fixed order in one precision")
    state = solver.prepare_state(positions, masses, leaf_size=int(leaf_size), max_order=int(order),
                                 theta=float(theta))
    return build_mutual_topology_from_state(state), state


# ------------------------------------------------------------------------ state
class MutualCapacityOverflow(RuntimeError):
    """A topology did not fit the capacities its state was built for
    (force.py:100-107); raised rather than truncating."""


class MutualCapacities(NamedTuple):
    """Pair-list and level-schedule capacities (force.py:110-155).  Device
    buffers here are sized to each topology, so capacities only bound it."""

    near: int
    far: int
    depth: int
    width: int


class MutualForceResult(NamedTuple):
    """Accelerations plus the near/far force split (force.py:82-97), input order."""

    accelerations: Any
    near_forces: Any
    far_forces: Any


@dataclass
class MutualFMMState:
    """Device state of the mutual lane (force.py:282-416) over a prepared FMM state."""

    prepared: Any
    near_a: torch.Tensor
    near_b: torch.Tensor
    self_leaves: torch.Tensor
    far_a: torch.Tensor
    far_b: torch.Tensor
    forward_permutation: torch.Tensor
    inverse_permutation: torch.Tensor
    softening: float
    G: float
    order: int
    use_pallas: bool = False
    near_chunk_size: Optional[int] = None
    pallas_interpret: bool = False
    num_particles_: int = 0
    num_near_pairs: int = 0
    num_far_pairs: int = 0
    topology_overflow: bool = False
    overflow_causes: int = 0
    flat: Any = None
    level_off: np.ndarray = None  # m2m level offsets (deepest first) for the cell rungs
    buffers: dict = field(default_factory=dict)
    near_items: Any = None  # work items over long near lists (nearsplit.P2PItems), None if every leaf fits one

    @property
    def num_particles(self) -> int:
        return int(self.num_particles_)

    @property
    def near_capacity(self) -> int:
        return int(self.near_a.shape[0])

    @property
    def far_capacity(self) -> int:
        return int(self.far_a.shape[0])


def build_mutual_state(topology: MutualTopology, *, softening: float, G: float = 1.0, use_pallas: bool = False,
                       near_chunk_size: Optional[int] = None, pallas_interpret: bool = False, index_dtype: Any = None,
                       caps: Optional[MutualCapacities] = None) -> MutualFMMState:
    """Device state for repeated evaluation (force.py:450-590).  ``use_pallas``,
    ``near_chunk_size``, ``pallas_interpret`` and ``index_dtype`` select the
    reference's JAX lanes and change nothing here."""
    from .flatlist import build_flat_list, pair_targets
    from .nearsplit import build_p2p_items

    del index_dtype
    prep = topology.prepared
    if prep is None:
        raise ValueError("the topology carries no prepared state; build it with build_mutual_topology")
    if caps is not None:
        depth = len(topology.level_nodes)
        width = max((int(u.shape[0]) for u in topology.level_nodes), default=0)
        if (topology.num_near_pairs > caps.near or topology.num_far_pairs > caps.far or depth > caps.depth
                or width > caps.width):
            raise MutualCapacityOverflow(
                f"topology overflows its capacities: near {topology.num_near_pairs}/{caps.near}, far "
                f"{topology.num_far_pairs}/{caps.far}, depth {depth}/{caps.depth}, width {width}/{caps.width}")
    dev = prep.device
    I = topology.num_internal

    def d32(a):
        return torch.as_tensor(np.ascontiguousarray(a, np.int32)).to(dev)

    K = prep.topology.num_far_pairs
    flat = build_flat_list(pair_targets(prep), prep.far_src[:K], prep.order, prep.dtype, dev) if K else None
    n = topology.num_particles
    st = MutualFMMState(
        prepared=prep, near_a=d32(topology.near_a - I), near_b=d32(topology.near_b - I),
        self_leaves=torch.arange(topology.num_leaves, dtype=torch.int32, device=dev), far_a=d32(topology.far_a),
        far_b=d32(topology.far_b), forward_permutation=d32(topology.forward_permutation),
        inverse_permutation=d32(topology.inverse_permutation), softening=float(softening), G=float(G),
        order=int(topology.order), use_pallas=bool(use_pallas), near_chunk_size=near_chunk_size,
        pallas_interpret=bool(pallas_interpret), num_particles_=n, num_near_pairs=topology.num_near_pairs,
        num_far_pairs=topology.num_far_pairs, flat=flat, level_off=np.asarray(prep.topology.m2m_off),
        near_items=build_p2p_items(np.asarray(prep.topology.near_off), prep.topology.leaf_size, prep.dtype, dev))
    st.buffers = {
        "far_acc": torch.empty((n, 3), dtype=prep.dtype, device=dev),
        "rung_sorted": torch.empty(n, dtype=torch.int32, device=dev),
        "node_rung": torch.empty(prep.topology.num_nodes, dtype=torch.int32, device=dev),
    }
    return st


# ------------------------------------------------------------------- evaluation
def _inputs(st: MutualFMMState, positions, masses):
    prep = st.prepared
    if isinstance(positions, torch.Tensor):
        kind = ("torch", positions.dtype, positions.device)
        x = positions.detach().to(prep.device, prep.dtype).contiguous()
        m = torch.as_tensor(masses).detach().to(prep.device, prep.dtype).contiguous()
    else:
        xn = np.asarray(positions)
        kind = ("numpy", xn.dtype, None)
        x = torch.as_tensor(np.ascontiguousarray(xn, np.float64)).to(prep.device, prep.dtype)
        m = torch.as_tensor(np.ascontiguousarray(masses, np.float64)).to(prep.device, prep.dtype)
    n = st.num_particles
    if tuple(x.shape) != (n, 3) or tuple(m.shape) != (n,):
        raise ValueError(f"positions/masses must have shapes ({n}, 3)/({n},) for this state")
    return x, m, kind


def _output(t: torch.Tensor, kind):
    if kind[0] == "numpy":
        dt = kind[1] if np.issubdtype(kind[1], np.floating) else np.float64
        return t.detach().cpu().numpy().astype(dt, copy=False)
    dt = kind[1] if kind[1] in (torch.float32, torch.float64) else t.dtype
    return t.to(device=kind[2], dtype=dt)


def _levels(st: MutualFMMState, rung, level_weights):
    """Level weights on the device and the particle / cell rungs (None when unweighted)."""
    if level_weights is None:
        return None
    if rung is None:
        raise ValueError("level_weights need a per-particle rung")
    prep = st.prepared
    topo = prep.topology
    s = nat.current_stream()
    lw = torch.as_tensor(np.asarray(level_weights, np.float64) if not isinstance(level_weights, torch.Tensor)
                         else level_weights.detach()).to(prep.device, prep.dtype).reshape(-1).contiguous()
    if not 1 <= lw.numel() <= MAX_LEVELS:
        raise ValueError(f"level_weights must have 1..{MAX_LEVELS} entries")
    r = torch.as_tensor(np.asarray(rung) if not isinstance(rung, torch.Tensor) else rung)
    r = r.to(prep.device, torch.int32).reshape(-1).contiguous()
    if r.numel() != topo.num_particles:
        raise ValueError(f"rung must have {topo.num_particles} entries")
    nat.call("jfm_mutual_rungs", s, topo.num_particles, topo.num_leaves, topo.leaf_size, prep.perm.data_ptr(),
             r.data_ptr(), st.buffers["rung_sorted"].data_ptr(), st.buffers["node_rung"].data_ptr())
    off = st.level_off
    for lev in range(len(off) - 1):  # deepest level first
        a, b = int(off[lev]), int(off[lev + 1])
        if b > a:
            nat.call("jfm_mutual_rungs_level", s, prep.m2m_nodes.data_ptr() + 4 * a, b - a, prep.left.data_ptr(),
                     prep.right.data_ptr(), st.buffers["node_rung"].data_ptr())
    return lw


def _tree_level(prep, nodes: torch.Tensor, off, op: str, rows: torch.Tensor, code: int, s) -> None:
    """One pass of M2M (deepest level first) or L2L (root first) over ``rows``."""
    for lev in range(len(off) - 1):
        a, b = int(off[lev]), int(off[lev + 1])
        if b <= a:
            continue
        if op == "m2m":
            nat.call("jfm_m2m_level", s, code, prep.order, nodes.data_ptr() + 4 * a, b - a, prep.left.data_ptr(),
                     prep.right.data_ptr(), prep.node_range.data_ptr(), rows.data_ptr(), prep.centers4.data_ptr())
        else:
            nat.call("jfm_l2l_level", s, code, prep.order, nodes.data_ptr() + 4 * a, b - a, prep.parent.data_ptr(),
                     prep.centers4.data_ptr(), rows.data_ptr())


def _far_m2l(st: MutualFMMState, lw, src: torch.Tensor, dst: torch.Tensor, code: int, s) -> None:
    """dst += M2L of src over both directions of every far pair (level-weighted when lw is given).
    On the symmetric list with symmetric weights this is also the transposed operator."""
    fl = st.flat
    if fl is None or not fl.num_pairs:
        return
    p = st.prepared.order
    c4 = st.prepared.centers4
    if lw is not None:
        nat.call("jfm_m2l_flat_weighted", s, code, p, fl.num_pairs, fl.pair_tgt.data_ptr(), fl.pair_src.data_ptr(),
                 st.buffers["node_rung"].data_ptr(), lw.data_ptr(), int(lw.numel()), src.data_ptr(), c4.data_ptr(),
                 dst.data_ptr(), fl.scratch.data_ptr(), fl.num_span, fl.span_tgt.data_ptr(), fl.span_w0.data_ptr(),
                 fl.span_w1.data_ptr())
    else:
        nat.call("jfm_m2l_flat", s, code, p, p, fl.num_pairs, fl.pair_tgt.data_ptr(), fl.pair_src.data_ptr(),
                 src.data_ptr(), c4.data_ptr(), dst.data_ptr(), fl.scratch.data_ptr(), fl.num_span,
                 fl.span_tgt.data_ptr(), fl.span_w0.data_ptr(), fl.span_w1.data_ptr())


def _expansions(st: MutualFMMState, x: torch.Tensor, m: torch.Tensor, lw, gathered: bool = False) -> torch.Tensor:
    """Gather (unless ``gathered``), P2M, M2M, (weighted) M2L, L2L at the live
    inputs; returns the locals after M2L alone (the full locals stay in the
    prepared state)."""
    prep = st.prepared
    topo = prep.topology
    s = nat.current_stream()
    code = nat.dtype_code(prep.dtype)
    n = topo.num_particles
    if not gathered:
        nat.call("jfm_gather_sorted", s, code, code, x.data_ptr(), m.data_ptr(), prep.perm.data_ptr(), n,
                 prep.pos4.data_ptr())
    nat.call("jfm_p2m", s, code, prep.order, prep.pos4.data_ptr(), n, topo.num_leaves, topo.leaf_size,
             topo.num_internal, prep.multipoles.data_ptr(), prep.centers4.data_ptr())
    _tree_level(prep, prep.m2m_nodes, topo.m2m_off, "m2m", prep.multipoles, code, s)
    prep.locals.zero_()
    _far_m2l(st, lw, prep.multipoles, prep.locals, code, s)
    lown = prep.locals.clone()
    _tree_level(prep, prep.l2l_nodes, topo.l2l_off, "l2l", prep.locals, code, s)
    return lown


def _evaluate(st: MutualFMMState, x: torch.Tensor, m: torch.Tensor, lw):
    """(acc, near_force, far_force) device tensors in input order."""
    prep = st.prepared
    topo = prep.topology
    s = nat.current_stream()
    code = nat.dtype_code(prep.dtype)
    n, L, ls, I = topo.num_particles, topo.num_leaves, topo.leaf_size, topo.num_internal
    nat.call("jfm_gather_sorted", s, code, code, x.data_ptr(), m.data_ptr(), prep.perm.data_ptr(), n,
             prep.pos4.data_ptr())
    near = torch.empty((n, 3), dtype=prep.dtype, device=prep.device)
    # the near field (reads only the gathered particles) runs on the P2P
    # stream of the device beside the tree passes and L2P on this one
    from .solver import FastMultipoleMethod

    main = torch.cuda.current_stream(prep.device)
    side = FastMultipoleMethod._side_stream(prep.device)
    side.wait_stream(main)
    with torch.cuda.stream(side):
        _near_field(st, lw, near, code, nat.current_stream())
    _expansions(st, x, m, lw, gathered=True)
    far_acc = st.buffers["far_acc"]
    nat.call("jfm_l2p", s, code, prep.order, prep.pos4.data_ptr(), n, ls, I, prep.locals.data_ptr(),
             prep.centers4.data_ptr(), st.G, prep.perm.data_ptr(), far_acc.data_ptr(), None, 0)
    main.wait_stream(side)
    acc = torch.empty((n, 3), dtype=prep.dtype, device=prep.device)
    far = torch.empty((n, 3), dtype=prep.dtype, device=prep.device)
    nat.call("jfm_mutual_combine", s, code, n, m.data_ptr(), far_acc.data_ptr(), near.data_ptr(), acc.data_ptr(),
             far.data_ptr())
    return acc, near, far


def _near_field(st: MutualFMMState, lw, near: torch.Tensor, code: int, s) -> None:
    """Rung-weighted mutual near field of the gathered particles into ``near`` (input order)."""
    prep = st.prepared
    topo = prep.topology
    n, L, ls = topo.num_particles, topo.num_leaves, topo.leaf_size
    weighted = lw is not None
    rs = st.buffers["rung_sorted"].data_ptr() if weighted else None
    lwp, nlev = (lw.data_ptr(), int(lw.numel())) if weighted else (None, 0)
    it = st.near_items
    if it is None:
        nat.call("jfm_mutual_p2p", s, code, prep.pos4.data_ptr(), n, L, ls, prep.near_off.data_ptr(),
                 prep.near_src.data_ptr(), rs, lwp, nlev, st.softening ** 2, st.G, prep.perm.data_ptr(),
                 near.data_ptr())
    else:
        nat.call("jfm_mutual_p2p_items", s, code, prep.pos4.data_ptr(), n, L, ls, it.num_items, it.leaf.data_ptr(),
                 it.s0.data_ptr(), it.s1.data_ptr(), it.part.data_ptr(), it.num_part, it.partials.data_ptr(),
                 it.num_split, it.split_leaf.data_ptr(), it.split_p0.data_ptr(), it.split_p1.data_ptr(),
                 prep.near_off.data_ptr(), prep.near_src.data_ptr(), rs, lwp, nlev, st.softening ** 2, st.G,
                 prep.perm.data_ptr(), near.data_ptr())


def _near_adjoint(st: MutualFMMState, lw, cot4, grad4, part, code: int, s) -> None:
    """Near-field part of the particle gradient into grad4 (written, not added);
    ``part``: partials of its own for the work items (None without items)."""
    prep = st.prepared
    topo = prep.topology
    n, L, ls = topo.num_particles, topo.num_leaves, topo.leaf_size
    eps2 = st.softening ** 2
    it = st.near_items
    if it is None and lw is not None:
        nat.call("jfm_p2p_adjoint_weighted", s, code, prep.pos4.data_ptr(), cot4.data_ptr(), n, L, ls,
                 prep.near_off.data_ptr(), prep.near_src.data_ptr(), st.buffers["rung_sorted"].data_ptr(),
                 lw.data_ptr(), int(lw.numel()), eps2, st.G, grad4.data_ptr())
    elif it is None:
        nat.call("jfm_p2p_adjoint", s, code, prep.pos4.data_ptr(), cot4.data_ptr(), n, L, ls,
                 prep.near_off.data_ptr(), prep.near_src.data_ptr(), eps2, st.G, grad4.data_ptr())
    else:  # long near lists cut into the forward's work items
        head = (s, code, prep.pos4.data_ptr(), cot4.data_ptr(), n, L, ls, it.num_items, it.leaf.data_ptr(),
                it.s0.data_ptr(), it.s1.data_ptr(), it.part.data_ptr(), it.num_part, part.data_ptr(),
                prep.near_off.data_ptr(), prep.near_src.data_ptr())
        if lw is not None:
            nat.call("jfm_p2p_adjoint_weighted_items", *head, st.buffers["rung_sorted"].data_ptr(), lw.data_ptr(),
                     int(lw.numel()), eps2, st.G, grad4.data_ptr())
        else:
            nat.call("jfm_p2p_adjoint_items", *head, eps2, st.G, grad4.data_ptr())
        nat.call("jfm_p2p_adjoint_reduce", s, code, cot4.data_ptr(), n, ls, it.num_split, it.split_leaf.data_ptr(),
                 it.split_p0.data_ptr(), it.split_p1.data_ptr(), part.data_ptr(), eps2, st.G, grad4.data_ptr())


def _vjp(st: MutualFMMState, x: torch.Tensor, m: torch.Tensor, lw, acc_bar: torch.Tensor):
    """(positions_bar, masses_bar) in input order for the cotangent of the mutual
    accelerations.  For m_i > 0 the acceleration is the level-weighted FMM
    acceleration, so the reverse sweep is the fixed-order adjoint
    (jaccpot_amd.autodiff) with every near and far pair carrying its weight;
    rows of zero mass are constant zero and pass no cotangent."""
    prep = st.prepared
    topo = prep.topology
    s = nat.current_stream()
    code = nat.dtype_code(prep.dtype)
    n, L, ls, I = topo.num_particles, topo.num_leaves, topo.leaf_size, topo.num_internal
    dt, dev = prep.dtype, prep.device
    nat.call("jfm_gather_sorted", s, code, code, x.data_ptr(), m.data_ptr(), prep.perm.data_ptr(), n,
             prep.pos4.data_ptr())
    ab = torch.where((m > 0)[:, None], acc_bar.to(dev, dt), torch.zeros((), dtype=dt, device=dev)).contiguous()
    pb = torch.zeros(n, dtype=dt, device=dev)
    cot4 = torch.empty((n, 4), dtype=dt, device=dev)
    nat.call("jfm_gather_sorted", s, code, code, ab.data_ptr(), pb.data_ptr(), prep.perm.data_ptr(), n,
             cot4.data_ptr())
    grad4 = torch.empty((n, 4), dtype=dt, device=dev)
    part = torch.empty_like(st.near_items.partials) if st.near_items is not None else None
    # the adjoint near field runs on the P2P stream of the device while this
    # stream recomputes the expansions and runs the tree adjoints; the L2P
    # part of the particle gradient goes to its own rows, added after the join
    from .solver import FastMultipoleMethod

    main = torch.cuda.current_stream(dev)
    side = FastMultipoleMethod._side_stream(dev)
    side.wait_stream(main)
    with torch.cuda.stream(side):
        _near_adjoint(st, lw, cot4, grad4, part, code, nat.current_stream())
    lown = _expansions(st, x, m, lw, gathered=True)
    grad_l4 = torch.zeros((n, 4), dtype=dt, device=dev)
    lbar = torch.zeros_like(prep.locals)
    nat.call("jfm_leaf_adjoint_a", s, code, prep.order, prep.pos4.data_ptr(), cot4.data_ptr(), n, ls, I,
             prep.locals.data_ptr(), prep.centers4.data_ptr(), st.G, grad_l4.data_ptr(), lbar.data_ptr())
    _tree_level(prep, prep.m2m_nodes, topo.m2m_off, "m2m", lbar, code, s)  # l2l^T
    mbar = torch.zeros_like(prep.multipoles)
    cbar4 = torch.zeros((topo.num_nodes, 4), dtype=dt, device=dev)
    if topo.m2l_targets.size:
        _far_m2l(st, lw, lbar, mbar, code, s)  # m2l^T on the symmetric list
        nt = int(topo.m2l_targets.size)
        if lw is not None:
            nat.call("jfm_center_adjoint_weighted", s, code, prep.order, prep.m2l_targets.data_ptr(), nt,
                     prep.far_off.data_ptr(), prep.far_src.data_ptr(), st.buffers["node_rung"].data_ptr(),
                     lw.data_ptr(), int(lw.numel()), prep.multipoles.data_ptr(), lown.data_ptr(), lbar.data_ptr(),
                     mbar.data_ptr(), prep.centers4.data_ptr(), cbar4.data_ptr())
        else:
            nat.call("jfm_center_adjoint", s, code, prep.order, prep.m2l_targets.data_ptr(), nt,
                     prep.far_off.data_ptr(), prep.far_src.data_ptr(), prep.multipoles.data_ptr(), lown.data_ptr(),
                     lbar.data_ptr(), mbar.data_ptr(), prep.centers4.data_ptr(), cbar4.data_ptr())
    _tree_level(prep, prep.l2l_nodes, topo.l2l_off, "l2l", mbar, code, s)  # m2m^T
    chain = torch.empty((L, 8), dtype=dt, device=dev)
    nat.call("jfm_com_chain", s, code, L, I, prep.parent.data_ptr(), prep.node_range.data_ptr(),
             prep.centers4.data_ptr(), cbar4.data_ptr(), chain.data_ptr())
    main.wait_stream(side)
    grad4.add_(grad_l4)
    gpos = torch.empty((n, 3), dtype=dt, device=dev)
    gmass = torch.empty((n,), dtype=dt, device=dev)
    nat.call("jfm_leaf_adjoint_b", s, code, prep.order, prep.pos4.data_ptr(), n, ls, I, mbar.data_ptr(),
             prep.centers4.data_ptr(), chain.data_ptr(), grad4.data_ptr(), prep.perm.data_ptr(), gpos.data_ptr(),
             gmass.data_ptr())
    return gpos, gmass


class _MutualForce(torch.autograd.Function):
    """torch.autograd over the mutual accelerations (the reference's jax.grad
    over mutual_weighted_accelerations, at fixed topology, rungs and weights)."""

    @staticmethod
    def forward(ctx, positions, masses, st, lw):
        x = positions.detach().contiguous()
        m = masses.detach().contiguous()
        acc, _, _ = _evaluate(st, x, m, lw)
        ctx.st, ctx.lw = st, lw
        # the rung buffers are shared by every evaluation of the state
        ctx.rungs = (st.buffers["rung_sorted"].clone(), st.buffers["node_rung"].clone()) if lw is not None else None
        ctx.save_for_backward(x, m)
        return acc

    @staticmethod
    def backward(ctx, gacc):
        x, m = ctx.saved_tensors
        if ctx.rungs is not None:
            ctx.st.buffers["rung_sorted"].copy_(ctx.rungs[0])
            ctx.st.buffers["node_rung"].copy_(ctx.rungs[1])
        gx, gm = _vjp(ctx.st, x, m, ctx.lw, gacc.contiguous())
        return gx, gm, None, None


def mutual_weighted_accelerations(state: MutualFMMState, positions, masses, *, rung=None, level_weights=None,
                                  return_parts: bool = False):
    """``sum_k level_weights[k] a_k`` in one traversal (force.py:985-1068); with
    ``level_weights=None`` the full acceleration.  Inputs and outputs in the
    caller's order; zero-mass particles get zero acceleration.  Torch inputs
    that require grad get gradients through the adjoint kernels (the
    accelerations; the near / far parts are returned detached)."""
    lw = _levels(state, rung, level_weights)
    grad = any(isinstance(t, torch.Tensor) and t.requires_grad for t in (positions, masses))
    if grad:
        prep = state.prepared
        kind = ("torch", positions.dtype, positions.device)
        x = positions.to(prep.device, prep.dtype)
        m = torch.as_tensor(masses).to(prep.device, prep.dtype)
        n = state.num_particles
        if tuple(x.shape) != (n, 3) or tuple(m.shape) != (n,):
            raise ValueError(f"positions/masses must have shapes ({n}, 3)/({n},) for this state")
        acc = _MutualForce.apply(x, m, state, lw)
        if return_parts:
            _, near, far = _evaluate(state, x.detach().contiguous(), m.detach().contiguous(), lw)
            return MutualForceResult(_output(acc, kind), _output(near, kind), _output(far, kind))
        return _output(acc, kind)
    x, m, kind = _inputs(state, positions, masses)
    acc, near, far = _evaluate(state, x, m, lw)
    if return_parts:
        return MutualForceResult(_output(acc, kind), _output(near, kind), _output(far, kind))
    return _output(acc, kind)


def mutual_accelerations(state: MutualFMMState, positions, masses, **kwargs):
    """Full mutual acceleration (force.py:1071-1093)."""
    return mutual_weighted_accelerations(state, positions, masses, **kwargs)


def mutual_level_accelerations(state: MutualFMMState, positions, masses, *, rung, level: int, k_max: int):
    """Level-``level`` acceleration alone (force.py:1096-1135)."""
    w = np.zeros(int(k_max) + 1)
    w[int(level)] = 1.0
    return mutual_weighted_accelerations(state, positions, masses, rung=rung, level_weights=w)


def mutual_near_field_forces(state: MutualFMMState, positions, masses, *, rung=None, level_weights=None):
    """Near-field forces alone (nearfield.py), input order."""
    return mutual_weighted_accelerations(state, positions, masses, rung=rung, level_weights=level_weights,
                                         return_parts=True).near_forces


def mutual_far_field_forces(state: MutualFMMState, positions, masses, *, rung=None, level_weights=None):
    """Far-field forces alone (farfield.py), input order."""
    return mutual_weighted_accelerations(state, positions, masses, rung=rung, level_weights=level_weights,
                                         return_parts=True).far_forces


# ------------------------------------------------------- block-step schedule
def n_sub(k_max: int) -> int:
    """Smallest sub-steps per base step, 2**k_max (force.py:685-698)."""
    return 1 << int(k_max)


def is_sync_boundary(s: int, k_max: int) -> bool:
    """Boundary s synchronises every rung (force.py:701-719)."""
    return int(s) == 0 or int(s) == (1 << int(k_max))


def active_level_floor(s: int, k_max: int) -> int:
    """Smallest level kicked at boundary s (force.py:722-745)."""
    s, k_max = int(s), int(k_max)
    if is_sync_boundary(s, k_max):
        return 0
    return k_max - ((s & -s).bit_length() - 1)


def level_weights_from_floor(active_floor: int, k_max: int, dt_max: float, *, half: float = 1.0,
                             dtype=np.float64) -> np.ndarray:
    """Kick weights half * dt_max / 2**k for k >= active_floor, else 0 (force.py:748-809)."""
    return np.asarray([(float(half) * float(dt_max) / float(1 << k)) if k >= int(active_floor) else 0.0
                       for k in range(int(k_max) + 1)], dtype=dtype)


def boundary_level_weights(s: int, k_max: int, dt_max: float, *, dtype=np.float64) -> np.ndarray:
    """Weights of boundary s: half at the synchronised ends (force.py:850-882)."""
    return level_weights_from_floor(active_level_floor(s, k_max), k_max, dt_max,
                                    half=0.5 if is_sync_boundary(s, k_max) else 1.0, dtype=dtype)


def boundary_weight_table(k_max: int, dt_max: float, *, dtype=np.float64) -> np.ndarray:
    """(n_sub + 1, k_max + 1) table of every boundary's weights (force.py:812-847)."""
    return np.stack([boundary_level_weights(s, k_max, dt_max, dtype=dtype) for s in range(n_sub(k_max) + 1)])


# ------------------------------------------------------------ block-step model
class BlockStepFMM:
    """Rung-aware, momentum-conserving force model (nornax_adapter.py:79-1212).

    ``basis`` must be ``"real"``; ``backend``, ``near_chunk_size``,
    ``pallas_interpret``, ``topology_backend`` and ``static_shapes`` select the
    reference's JAX lanes and change nothing here (the topology is built by
    this package's builder, on the GPU for large N).
    """

    def __init__(self, softening: float, k_max: int, theta: float = 0.6, max_order: int = 4, G: float = 1.0,
                 basis: str = "real", backend: str = "jax", leaf_size: int = 64,
                 near_chunk_size: Optional[int] = None, pallas_interpret: bool = False,
                 topology_backend: str = "host", static_shapes: bool = False,
                 caps: Optional[MutualCapacities] = None, validate_rung: bool = True):
        if basis != "real":
            raise ValueError(f"BlockStepFMM supports basis='real' only; got {basis!r}")
        if backend not in ("jax", "pallas"):
            raise ValueError(f"backend must be 'jax' or 'pallas'; got {backend!r}")
        if int(k_max) < 0:
            raise ValueError(f"k_max must be >= 0; got {k_max!r}")
        if int(k_max) + 1 > MAX_LEVELS:
            raise ValueError(f"k_max must be below {MAX_LEVELS}")
        if topology_backend not in ("host", "device"):
            raise ValueError(f"topology_backend must be 'host' or 'device'; got {topology_backend!r}")
        if not 1 <= int(leaf_size) <= 256:
            raise ValueError("leaf_size must be in 1..256 on this lane")
        self.softening = float(softening)
        self.k_max = int(k_max)
        self.theta = float(theta)
        self.max_order = int(max_order)
        self.G = float(G)
        self.leaf_size = int(leaf_size)
        self.backend = backend
        self.topology_backend = topology_backend
        self.static_shapes = bool(static_shapes) or caps is not None
        self._caps = caps
        self.validate_rung = bool(validate_rung)
        self._state: Optional[MutualFMMState] = None
        self._topology: Optional[MutualTopology] = None

    @property
    def state(self) -> Optional[MutualFMMState]:
        return self._state

    @property
    def capacities(self) -> Optional[MutualCapacities]:
        return self._caps

    def prepare(self, positions, masses) -> MutualFMMState:
        """Build the frozen topology for these inputs and cache its state."""
        topo, _ = build_mutual_topology(positions, masses, theta=self.theta, order=self.max_order,
                                        leaf_size=self.leaf_size)
        if self.static_shapes and self._caps is None:
            width = max((int(u.shape[0]) for u in topo.level_nodes), default=0)
            self._caps = MutualCapacities(near=topo.num_near_pairs, far=topo.num_far_pairs,
                                          depth=len(topo.level_nodes), width=width)
        self._topology = topo
        self._state = build_mutual_state(topo, softening=self.softening, G=self.G, caps=self._caps)
        return self._state

    rebuild_state = prepare

    def refresh(self, positions, masses) -> MutualFMMState:
        return self.prepare(positions, masses)

    def _require_state(self, positions, masses) -> MutualFMMState:
        if self._state is None:
            return self.prepare(positions, masses)
        return self._state

    def _validate_rung(self, rung):
        if self.validate_rung:
            r = rung.detach().cpu().numpy() if isinstance(rung, torch.Tensor) else np.asarray(rung)
            if r.size and (r.min() < 0 or r.max() > self.k_max):
                raise ValueError(f"rungs must lie in [0, k_max={self.k_max}]")
        return rung

    def weighted_accelerations(self, positions, masses, *, rung, level_weights):
        st = self._require_state(positions, masses)
        return mutual_weighted_accelerations(st, positions, masses, rung=self._validate_rung(rung),
                                             level_weights=level_weights)

    def boundary_weights(self, s: int, dt_max: float) -> np.ndarray:
        return boundary_level_weights(s, self.k_max, dt_max)

    def level_accelerations(self, positions, masses, *, rung, level: int, args: object = None):
        del args
        st = self._require_state(positions, masses)
        rung = self._validate_rung(rung)
        if not 0 <= int(level) <= self.k_max:
            raise ValueError(f"level must lie in [0, k_max={self.k_max}]; got {level!r}")
        return mutual_level_accelerations(st, positions, masses, rung=rung, level=int(level), k_max=self.k_max)

    def total_accelerations(self, positions, masses, *, rung=None, args: object = None):
        del args, rung
        return mutual_weighted_accelerations(self._require_state(positions, masses), positions, masses)

    def boundary_kick(self, positions, velocities, masses, *, rung, active_floor=None, dt_max=None, half=1.0,
                      level_weights=None, args: object = None):
        """velocities + sum_k w_k a_k for one sub-step boundary, one traversal
        (nornax_adapter.py:950-1047)."""
        del args
        st = self._require_state(positions, masses)
        rung = self._validate_rung(rung)
        if level_weights is None:
            if dt_max is None or active_floor is None:
                raise ValueError("boundary_kick needs either level_weights, or both active_floor and dt_max")
            level_weights = level_weights_from_floor(active_floor, self.k_max, dt_max, half=half)
        else:
            lw = np.asarray(level_weights.detach().cpu() if isinstance(level_weights, torch.Tensor)
                            else level_weights, np.float64)
            if lw.shape[-1] != self.k_max + 1:
                raise ValueError(f"level_weights must have {self.k_max + 1} entries for k_max={self.k_max}; "
                                 f"got shape {tuple(lw.shape)}")
            level_weights = lw
        dv = mutual_weighted_accelerations(st, positions, masses, rung=rung, level_weights=level_weights)
        return velocities + dv

    def boundary_kick_at(self, positions, velocities, masses, *, rung, s: int, dt_max: float, args: object = None):
        return self.boundary_kick(positions, velocities, masses, rung=rung,
                                  active_floor=active_level_floor(int(s), self.k_max), dt_max=dt_max,
                                  half=0.5 if is_sync_boundary(int(s), self.k_max) else 1.0, args=args)

    def advance_base_step(self, positions, velocities, masses, *, rung, dt_max: float,
                          scan_boundaries: bool = False):
        """One base step of the block-step KDK palindrome (nornax_adapter.py:1099-1212):
        a kick at every boundary s = 0..n_sub (half at the ends), a drift of
        dt_max / n_sub between them, rungs fixed; returns (positions,
        velocities, full acceleration at the final positions)."""
        del scan_boundaries
        st = self._require_state(positions, masses)
        rung = self._validate_rung(rung)
        steps = n_sub(self.k_max)
        dt_min = float(dt_max) / steps
        pos, vel = positions, velocities
        for s in range(steps + 1):
            w = boundary_level_weights(s, self.k_max, float(dt_max))
            vel = vel + mutual_weighted_accelerations(st, pos, masses, rung=rung, level_weights=w)
            if s < steps:
                pos = pos + dt_min * vel
        return pos, vel, mutual_weighted_accelerations(st, pos, masses)
