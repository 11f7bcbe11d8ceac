"""``FastMultipoleMethod`` -- the reference's public facade, on libjfm.

Mirrors jaccpot/solver.py:540-1356 for the fixed-topology, real-basis path:

* ``prepare_state(positions, masses, *, bounds, leaf_size, max_order, theta)``
  (solver.py:1012-1084): host topology (tree + far/near lists), then on the
  GPU the gather into tree order, leaf P2M, M2M by level, M2L over the far
  list and the L2L cascade; returns an :class:`FMMPreparedState`.
* ``evaluate_prepared_state(state, *, target_indices, return_potential)``
  (solver.py:1292-1356): fused L2P + P2P on the GPU, results in input order.
* ``compute_accelerations`` = prepare + evaluate (fmm_evaluate.py:289-322).

Errors follow the reference: ``ValueError`` for bad configuration
(solver.py:709-722), ``NotImplementedError`` for unsupported combinations.
Every numeric step runs in libjfm's HIP kernels; when the library or a GPU is
missing the constructor raises instead of falling back.
"""

from __future__ import annotations

import weakref
from dataclasses import dataclass, field
from typing import Any, Optional, Sequence

import numpy as np
import torch

from . import _native as nat
from .flatlist import build_flat_list, launch_flat, pair_targets, prefers_flat
from .nearsplit import build_p2p_items
from .dtopology import build_topology_device
from .topology import Topology, build_topology

_PRESETS = ("fast", "balanced", "accurate", "large_n_gpu")
_SIDE_STREAMS: dict = {}  # device -> the overlapped step's P2P stream (FastMultipoleMethod._side_stream)


def _require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("jaccpot_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    nat.load()
    return torch.device("cuda", torch.cuda.current_device())


@dataclass
class FMMPreparedState:
    """Frozen topology + device buffers of one prepared evaluation.

    Counterpart of FMMPreparedState (runtime/fmm_state.py:802-1073): host
    topology, tree-order particles, node centres, multipoles and locals.
    Host-built and not traceable, like the reference's (solver.py:1068-1070).
    """

    topology: Topology
    order: int
    dtype: torch.dtype
    G: float
    softening: float
    theta: float
    device: torch.device
    input_kind: str  # "numpy" or "torch"
    input_dtype: Any
    input_device: Any
    cs: int
    perm: torch.Tensor
    pos4: torch.Tensor
    node_range: torch.Tensor
    left: torch.Tensor
    right: torch.Tensor
    parent: torch.Tensor
    m2m_nodes: torch.Tensor
    l2l_nodes: torch.Tensor
    m2l_targets: torch.Tensor
    far_off: torch.Tensor
    far_src: torch.Tensor
    near_off: torch.Tensor
    near_src: torch.Tensor
    centers4: torch.Tensor
    multipoles: torch.Tensor
    locals: torch.Tensor
    compensated: bool = True
    extra: dict = field(default_factory=dict)

    @property
    def num_particles(self) -> int:
        return self.topology.num_particles


class FastMultipoleMethod:
    """MI355X-native FMM facade with the reference's constructor surface.

    Parameters mirror jaccpot/solver.py:664-692.  Only the real basis and the
    fixed radix topology are implemented; ``m2l_impl`` accepts ``None`` or
    ``"rot_scale"`` (the reference's real-basis default, solver.py:723-726).
    ``precision`` picks float32 or float64 kernels; ``None`` follows the input
    dtype.  ``compensated`` (default on) selects blocked Kahan accumulation in
    the float32 P2P sums.  ``mixed_precision`` (float32 only) runs the M2M and
    L2L of the top three tree levels in float64 on shadow rows.
    ``adaptive_order=True`` with a trained ``order_model``
    (:class:`jaccpot_amd.learned.OrderModel`) evaluates each far pair at the
    smallest gear of ``p_gears`` whose predicted force error (relative
    truncation error times m_s / d^2) is at most ``10**gear_tolerance`` times
    the largest such error of the top gear among the target's pairs (the
    largest gear otherwise); ``max_order`` must be the largest gear.  Gears
    are fixed in ``prepare_state``.  ``tree_build`` picks where ``prepare_state``
    builds the topology: ``"device"`` (HIP kernels, default) or ``"host"``
    (C++ threads); both give the same arrays bit for bit.  ``split_near``
    (default on) cuts long near lists into P2P work items of at most 8192
    source slots (``nearsplit``), so a few leaves with huge near lists do not
    run alone at the end of the P2P launch.  ``m2l_engine`` picks where the
    float32 M2L applies its constant B_U rotation blocks: ``"valu"`` (unrolled
    FMAs with the constants folded in; the default, faster on MI355X) or
    ``"mfma"`` (matrix cores, ``v_mfma_f32_4x4x1_16b_f32``); float64 always
    uses the VALU.  ``reuse_topology`` with
    ``rebuild_every`` (solver.py:592-597) keeps the tree and the far/near lists
    of the last ``prepare_state`` for up to ``rebuild_every - 1`` further calls
    whose sorted Morton codes are unchanged (runtime/fmm_prepare.py:797-802):
    the particle order, centres, expansions and locals are recomputed, the
    interaction lists are not.
    """

    # Overlapped step: run the P2M on the current stream before the P2P is
    # forked onto its own stream (alone, at the full width of the GPU) instead
    # of beside the P2P's workgroups.  Same results either way.
    _p2m_before_near = True

    def __init__(
        self,
        *,
        preset: str = "fast",
        basis: str = "real",
        m2l_impl: Optional[str] = None,
        adaptive_order: bool = False,
        p_gears: Optional[Sequence[int]] = None,
        use_pallas: Optional[bool] = None,
        reuse_topology: bool = False,
        rebuild_every: int = 1,
        mac_force_scale_mode: str = "prev",
        mac_force_scale_prepass_theta: Optional[float] = None,
        mac_force_scale_fb_inflation: float = 1.0,
        adaptive_error_model: str = "tail_proxy",
        adaptive_eps: Optional[float] = None,
        dehnen_geometry_mode: str = "com",
        mac_theta_max: float = 1.0,
        theta: float = 0.6,
        G: float = 1.0,
        softening: float = 1e-3,
        precision: Optional[str] = None,
        working_dtype: Any = None,
        advanced: Any = None,
        compensated: bool = True,
        mixed_precision: bool = False,
        order_model: Any = None,
        gear_tolerance: float = -2.5,
        gear_budget: str = "node",
        order_judge_targets: int = 1024,
        order_judge_factor: float = 2.0,
        order_judge_seed: int = 0,
        num_threads: int = 0,
        tree_build: str = "device",
        split_near: bool = True,
        m2l_engine: str = "valu",
        **legacy_kwargs: Any,
    ):
        from .config import resolve_constructor_options

        opts = resolve_constructor_options(
            adaptive_order=bool(adaptive_order), use_pallas=use_pallas, mac_force_scale_mode=mac_force_scale_mode,
            mac_force_scale_prepass_theta=mac_force_scale_prepass_theta,
            mac_force_scale_fb_inflation=mac_force_scale_fb_inflation, adaptive_error_model=adaptive_error_model,
            adaptive_eps=adaptive_eps, dehnen_geometry_mode=dehnen_geometry_mode, mac_theta_max=mac_theta_max,
            advanced=advanced, legacy_kwargs=legacy_kwargs)
        basis = opts["renamed"].get("basis", basis)
        preset = opts["renamed"].get("preset", preset)
        preset_norm = str(getattr(preset, "value", preset)).strip().lower()
        if preset_norm not in _PRESETS:
            raise ValueError(f"unknown preset {preset!r}; expected one of {_PRESETS}")
        if str(getattr(basis, "value", basis)).strip().lower() != "real":
            raise NotImplementedError("only basis='real' (the Dehnen real basis) is implemented")
        if m2l_impl not in (None, "rot_scale"):
            raise NotImplementedError(f"m2l_impl={m2l_impl!r} is not implemented (use 'rot_scale')")
        if adaptive_order:
            if p_gears is None and order_model is None:
                raise ValueError("adaptive_order=True requires non-empty p_gears")
            gears = tuple(int(g) for g in (p_gears if p_gears is not None else order_model.gears))
            if len(gears) != 3 or list(gears) != sorted(set(gears)) or not (1 <= gears[0] and gears[-1] <= 8):
                raise ValueError(f"p_gears must be three increasing orders in 1..8, got {gears}")
            if order_model is not None and gears != tuple(order_model.gears):
                raise ValueError(f"p_gears {gears} do not match the order model's gears {order_model.gears}")
        elif p_gears is not None:
            raise ValueError("p_gears is only meaningful with adaptive_order=True")
        if not (theta > 0.0):
            raise ValueError("theta must be > 0")
        if not (softening >= 0.0):
            raise ValueError("softening must be >= 0")
        if tree_build not in ("device", "host"):
            raise ValueError("tree_build must be 'device' or 'host'")
        if int(rebuild_every) <= 0:
            raise ValueError("rebuild_every must be positive")
        if m2l_engine not in ("valu", "mfma"):
            raise ValueError("m2l_engine must be 'valu' or 'mfma'")
        if int(order_judge_targets) < 0 or not (float(order_judge_factor) >= 1.0):
            raise ValueError("order_judge_targets must be >= 0 and order_judge_factor >= 1")
        dtype = None
        if precision is not None:
            p = str(precision).strip().lower()
            if p not in ("fp32", "fp64"):
                raise ValueError("precision must be one of ('fp32', 'fp64')")
            dtype = torch.float32 if p == "fp32" else torch.float64
        if working_dtype is not None:
            wd = torch.float64 if np.dtype(str(working_dtype).replace("torch.", "")) == np.float64 else torch.float32
            if dtype is not None and wd != dtype:
                raise ValueError("precision conflicts with explicit working_dtype; use only one or ensure they match")
            dtype = wd
        if preset_norm == "large_n_gpu" and dtype is None:
            dtype = torch.float32
        self.preset = preset_norm
        self.basis = "real"
        self.options = opts
        self.use_pallas = use_pallas
        self.advanced = advanced
        self.theta = float(theta)
        self.G = float(G)
        self.softening = float(softening)
        self.working_dtype = dtype
        self.compensated = bool(compensated)
        self.mixed_precision = bool(mixed_precision)
        self.adaptive_order = bool(adaptive_order)
        self.gear_tolerance = float(gear_tolerance)
        if gear_budget not in ("node", "leaf"):
            raise ValueError(f"gear_budget must be 'node' or 'leaf', got {gear_budget!r}")
        self.gear_budget = gear_budget
        self.order_judge_targets = int(order_judge_targets)
        self.order_judge_factor = float(order_judge_factor)
        self.order_judge_seed = int(order_judge_seed)
        self.num_threads = int(num_threads)
        self.tree_build = tree_build
        self.split_near = bool(split_near)
        self.m2l_engine = m2l_engine
        self.reuse_topology = bool(reuse_topology)
        self.rebuild_every = int(rebuild_every)
        self._reuse: Optional[dict] = None
        self._last_inputs: Optional[tuple] = None
        self.device = _require_gpu()
        self._last_state: Optional[FMMPreparedState] = None
        if self.adaptive_order and order_model is None:
            order_model = self._shipped_order_model(gears)
        self.order_model = order_model
        self.p_gears = tuple(order_model.gears) if order_model is not None else None

    @staticmethod
    def _shipped_order_model(gears):
        """The order model trained for ``gears`` that ships with the package."""
        from .train_order import load, weights_path

        path = weights_path(gears)
        if not path.exists():
            raise NotImplementedError(
                f"no trained order model ships for p_gears={gears}; train one with "
                f"python -m jaccpot_amd.train_order --gears {','.join(map(str, gears))} and pass it as order_model")
        return load(path)

    # ------------------------------------------------------------------ utils
    def _as_inputs(self, positions, masses):
        if isinstance(positions, torch.Tensor):
            kind, in_dtype, in_dev = "torch", positions.dtype, positions.device
            x_host = positions.detach().to("cpu", torch.float64).numpy()
            m_host = torch.as_tensor(masses).detach().to("cpu", torch.float64).numpy()
        else:
            x_np = np.asarray(positions)
            kind, in_dtype, in_dev = "numpy", x_np.dtype, None
            x_host = np.asarray(x_np, np.float64)
            m_host = np.asarray(masses, np.float64)
        if x_host.ndim != 2 or x_host.shape[1] != 3:
            raise ValueError("positions must have shape (N, 3)")
        if m_host.shape != (x_host.shape[0],):
            raise ValueError("masses must have shape (N,) matching positions")
        if x_host.shape[0] < 1:
            raise ValueError("need at least one particle")
        return x_host, m_host, kind, in_dtype, in_dev

    def _work_dtype(self, in_dtype) -> torch.dtype:
        if self.working_dtype is not None:
            return self.working_dtype
        if in_dtype in (torch.float64, np.float64) or str(in_dtype) in ("float64", "torch.float64"):
            return torch.float64
        return torch.float32

    # -------------------------------------------------------------- prepare
    def prepare_state(self, positions, masses, *, bounds=None, leaf_size: int = 16, max_order: int = 4,
                      theta: Optional[float] = None, cache_policy: str = "auto", force_scale_nodes=None,
                      runtime_overrides_override=None, fused_device_mode: bool = False,
                      **extra) -> FMMPreparedState:
        """Build the topology and run the upward + downward passes (solver.py:1012).

        ``cache_policy``, ``runtime_overrides_override`` and
        ``fused_device_mode`` choose how the reference caches and lays out its
        JAX state; they do not change the result and have no effect here.
        With ``adaptive_order`` the learned gears are chosen here and judged
        against the plain top-gear FMM on ``order_judge_targets`` seeded
        targets (``jaccpot_amd.judge``); the verdict is in
        ``state.extra["order_judge"]``."""
        from .config import resolve_prepare_options

        resolve_prepare_options(adaptive_order=self.adaptive_order, cache_policy=cache_policy,
                                force_scale_nodes=force_scale_nodes,
                                runtime_overrides_override=runtime_overrides_override,
                                fused_device_mode=fused_device_mode, extra=extra)
        p = int(max_order)
        if p < 1 or p > nat.MAX_ORDER:
            raise ValueError(f"max_order must be in [1, {nat.MAX_ORDER}]")
        if int(leaf_size) < 1 or int(leaf_size) > 256:
            raise ValueError("leaf_size must be in [1, 256]")
        th = self.theta if theta is None else float(theta)
        if not (th > 0.0):
            raise ValueError("theta must be > 0")
        x_host, m_host, kind, in_dtype, in_dev = self._as_inputs(positions, masses)
        if not np.all(np.isfinite(x_host)) or not np.all(np.isfinite(m_host)):
            raise ValueError("positions and masses must be finite")
        dtype = self._work_dtype(in_dtype)
        dev = self.device
        xd = torch.as_tensor(x_host, device=dev)
        md = torch.as_tensor(m_host, device=dev)
        key = (x_host.shape[0], int(leaf_size), p, th, dtype, None if bounds is None else
               tuple(np.concatenate([np.asarray(bounds[0], float), np.asarray(bounds[1], float)]).tolist()))
        reused = self._try_reuse(key, xd, md, bounds, kind, in_dtype, in_dev)
        if reused is not None:
            return reused
        if self.tree_build == "device":
            topo = build_topology_device(xd, md, leaf_size=int(leaf_size), theta=th, bounds=bounds)
        else:
            topo = build_topology(x_host, m_host, leaf_size=int(leaf_size), theta=th, bounds=bounds,
                                  num_threads=self.num_threads)
        on_dev = topo.extra.get("device", {})
        state = self._state_from_topology(topo, p, dtype, th, kind, in_dtype, in_dev,
                                          mixed=self.mixed_precision and dtype == torch.float32,
                                          split_near=self.split_near, G=self.G, softening=self.softening,
                                          compensated=self.compensated)
        if self.adaptive_order:
            self._prepare_gears(state, xd, md)
        self.refresh_prepared_state(state, xd, md)
        if self.adaptive_order and self.order_judge_targets > 0 and state.extra.get("gears") is not None:
            from .judge import judge_learned_orders

            judge_learned_orders(self, state, xd, md, num_targets=self.order_judge_targets,
                                 factor=self.order_judge_factor, seed=self.order_judge_seed)
        self._last_state = state
        if self.reuse_topology:
            codes = on_dev.get("sorted_codes")
            if codes is None:
                from .dtopology import sorted_morton

                codes, _ = sorted_morton(xd, bounds)
            self._reuse = {"key": key, "codes": codes, "state": state, "count": 0}
        return state

    def _state_from_topology(self, topo: Topology, p: int, dtype: torch.dtype, th: float, kind: str, in_dtype,
                             in_dev, *, mixed: bool, split_near: bool, G: float, softening: float,
                             compensated: bool) -> FMMPreparedState:
        """Device buffers and schedules of a prepared state for ``topo`` (expansions zero)."""
        dev = self.device
        on_dev = topo.extra.get("device", {})
        cs = nat.coeff_stride(p, dtype)

        def i32(name):
            d = on_dev.get(name)
            if d is not None and d.numel() > 0:
                return d
            a = np.asarray(getattr(topo, name))
            return torch.as_tensor(np.ascontiguousarray(a if a.size else np.zeros(1), np.int32)).to(dev)

        nodes = topo.num_nodes
        state = FMMPreparedState(
            topology=topo, order=p, dtype=dtype, G=G, softening=softening, theta=th, device=dev,
            input_kind=kind, input_dtype=in_dtype, input_device=in_dev, cs=cs,
            perm=i32("perm"),
            pos4=torch.empty((topo.num_particles, 4), dtype=dtype, device=dev),
            node_range=i32("node_range"), left=i32("left"), right=i32("right"), parent=i32("parent"),
            m2m_nodes=i32("m2m_nodes"), l2l_nodes=i32("l2l_nodes"), m2l_targets=i32("m2l_targets"),
            far_off=i32("far_off"), far_src=i32("far_src"), near_off=i32("near_off"), near_src=i32("near_src"),
            centers4=torch.zeros((nodes, 4), dtype=dtype, device=dev),
            multipoles=torch.zeros((nodes, cs), dtype=dtype, device=dev),
            locals=torch.zeros((nodes, cs), dtype=dtype, device=dev),
            compensated=compensated,
        )
        if mixed:
            state.extra["mixed"] = self._mixed_schedule(topo, p, dev)
        if prefers_flat(topo.num_far_pairs, int(topo.m2l_targets.size)):
            state.extra["flat"] = build_flat_list(pair_targets(state), state.far_src[: topo.num_far_pairs], p, dtype,
                                                  dev)
        if split_near:
            state.extra["p2p_items"] = build_p2p_items(topo.near_off, topo.leaf_size, dtype, dev)
        return state

    def save_prepared_state(self, state: FMMPreparedState, path) -> None:
        """Write a prepared state to ``path`` (.npz: topology, tree-order particles,
        centres, expansions, per-pair gears; no pickled objects)."""
        from .persist import save_prepared_state

        save_prepared_state(state, path)

    def load_prepared_state(self, path) -> FMMPreparedState:
        """Read a state written by ``save_prepared_state`` onto this solver's device;
        it evaluates exactly as the saved one and can be refreshed like it."""
        from .persist import load_prepared_state

        return load_prepared_state(self, path)

    def _try_reuse(self, key, xd, md, bounds, kind, in_dtype, in_dev) -> Optional[FMMPreparedState]:
        """A state on the cached tree and lists when the new sorted Morton codes
        equal the cached ones and the reuse budget allows (fmm_prepare.py:797-802)."""
        import copy
        import dataclasses

        from .dtopology import DeviceTopology, _Lazy, sorted_morton

        r = self._reuse
        if not self.reuse_topology or r is None or r["key"] != key or r["count"] >= self.rebuild_every - 1:
            return None
        codes, perm = sorted_morton(xd, bounds)
        if not torch.equal(codes, r["codes"]):
            return None
        old = r["state"]
        topo = copy.copy(old.topology)
        topo.perm = _Lazy(perm) if isinstance(topo, DeviceTopology) else perm.cpu().numpy()
        topo.extra = dict(topo.extra)
        topo.extra["device"] = dict(topo.extra.get("device", {}), perm=perm)
        state = dataclasses.replace(
            old, topology=topo, perm=perm, input_kind=kind, input_dtype=in_dtype, input_device=in_dev,
            pos4=torch.empty_like(old.pos4), centers4=torch.zeros_like(old.centers4),
            multipoles=torch.zeros_like(old.multipoles), locals=torch.zeros_like(old.locals), extra=dict(old.extra))
        if "mixed" in state.extra:
            state.extra["mixed"] = self._mixed_schedule(topo, state.order, state.device)
        self.refresh_prepared_state(state, xd, md)
        r["count"] += 1
        r["state"] = state
        self._last_state = state
        return state

    def _prepare_gears(self, state: FMMPreparedState, xd: torch.Tensor, md: torch.Tensor) -> None:
        """Tag every far pair with a gear from the order model (prepare time, like the reference's traversal)."""
        from .learned import build_gear_schedule, pair_features

        if state.dtype != torch.float32:
            raise NotImplementedError("learned per-pair orders run on the float32 path")
        gears = tuple(self.order_model.gears)
        if state.order != max(gears):
            raise ValueError(f"max_order must equal the largest gear {max(gears)}")
        s = nat.current_stream()
        code = nat.dtype_code(state.dtype)
        nat.call("jfm_gather_sorted", s, nat.dtype_code(xd.dtype), code, xd.data_ptr(), md.to(xd.dtype).data_ptr(),
                 state.perm.data_ptr(), state.topology.num_particles, state.pos4.data_ptr())
        self._upward(state, s, code)
        K = state.topology.num_far_pairs
        if K == 0:
            return
        X = pair_features(state)
        g = self.order_model.select_gears(X, self.gear_tolerance, state, budget=self.gear_budget)
        state.extra["gears"] = build_gear_schedule(state, g, gears)
        state.extra["gear_of_pair"] = g.cpu().numpy()

    def refresh_prepared_state(self, state: FMMPreparedState, positions: torch.Tensor, masses: torch.Tensor) -> None:
        """Re-run gather + upward + downward at the state's frozen topology.

        ``positions``/``masses`` are device tensors in input order (float32 or
        float64).  The fixed-topology re-evaluation the reference performs in
        its differentiable path (fmm_evaluate.py:542-731).
        """
        topo = state.topology
        s = nat.current_stream()
        code = nat.dtype_code(state.dtype)
        src_code = nat.dtype_code(positions.dtype)
        positions = positions.contiguous()
        masses = masses.to(positions.dtype).contiguous()
        nat.call("jfm_gather_sorted", s, src_code, code, positions.data_ptr(), masses.data_ptr(),
                 state.perm.data_ptr(), topo.num_particles, state.pos4.data_ptr())
        self._upward(state, s, code)
        self._downward(state, s, code)

    @staticmethod
    def _mixed_schedule(topo: Topology, p: int, dev) -> dict:
        """Shadow rows and level tables for the float64 top three levels."""
        I = topo.num_internal
        depth = topo.depth
        top = np.nonzero(depth[:I] <= 2)[0].astype(np.int32)
        shadow = np.full(topo.num_nodes, -1, np.int32)
        shadow[top] = np.arange(top.size, dtype=np.int32)
        m2m_levels = [top[depth[top] == d] for d in (2, 1, 0)]
        m2m_nodes = np.concatenate(m2m_levels) if top.size else np.zeros(0, np.int32)
        m2m_off = np.concatenate([[0], np.cumsum([len(v) for v in m2m_levels])]).astype(np.int32)
        l2l_levels = [np.nonzero(depth == d)[0].astype(np.int32) for d in (1, 2, 3)]
        l2l_nodes = np.concatenate(l2l_levels)
        l2l_off = np.concatenate([[0], np.cumsum([len(v) for v in l2l_levels])]).astype(np.int32)

        def dev_i32(a):
            return torch.as_tensor(np.ascontiguousarray(a if a.size else np.zeros(1, np.int32), np.int32)).to(dev)

        cs64 = nat.coeff_stride(p, torch.float64)
        ntop = max(int(top.size), 1)
        return {
            "top": dev_i32(top), "ntop": int(top.size), "shadow": dev_i32(shadow),
            "m2m_nodes": dev_i32(m2m_nodes), "m2m_off": dev_i32(m2m_off), "m2m_off_host": m2m_off,
            "l2l_nodes": dev_i32(l2l_nodes), "l2l_off": dev_i32(l2l_off), "l2l_off_host": l2l_off,
            "mult64": torch.zeros((ntop, cs64), dtype=torch.float64, device=dev),
            "centers64": torch.zeros((ntop, 4), dtype=torch.float64, device=dev),
            "loc64": torch.zeros((ntop, cs64), dtype=torch.float64, device=dev),
        }

    @staticmethod
    def _p2m(state: FMMPreparedState, s, code) -> None:
        topo = state.topology
        nat.call("jfm_p2m", s, code, state.order, state.pos4.data_ptr(), topo.num_particles, topo.num_leaves,
                 topo.leaf_size, topo.num_internal, state.multipoles.data_ptr(), state.centers4.data_ptr())

    def _upward(self, state: FMMPreparedState, s, code, p2m: bool = True) -> None:
        topo = state.topology
        mixed = state.extra.get("mixed")
        if p2m:
            self._p2m(state, s, code)
        off = topo.m2m_off
        base = state.m2m_nodes.data_ptr()
        D = topo.max_depth
        for lev in range(len(off) - 1):
            a, b = int(off[lev]), int(off[lev + 1])
            if mixed is not None and D - lev <= 2:
                continue  # float64 top levels below
            if b > a:
                nat.call("jfm_m2m_level", s, code, state.order, base + 4 * a, b - a, state.left.data_ptr(),
                         state.right.data_ptr(), state.node_range.data_ptr(), state.multipoles.data_ptr(),
                         state.centers4.data_ptr())
        if mixed is not None and mixed["ntop"]:
            oh = mixed["m2m_off_host"]
            nat.call("jfm_m2m_top", s, state.order, mixed["m2m_nodes"].data_ptr(), mixed["m2m_off"].data_ptr(),
                     len(oh) - 1, nat.ptr(oh), state.left.data_ptr(), state.right.data_ptr(),
                     state.node_range.data_ptr(), mixed["shadow"].data_ptr(), state.multipoles.data_ptr(),
                     state.centers4.data_ptr(), mixed["mult64"].data_ptr(), mixed["centers64"].data_ptr())

    def _downward(self, state: FMMPreparedState, s, code, zeroed: bool = False) -> None:
        if not zeroed:
            state.locals.zero_()
        self._m2l(state, s, code)
        self._l2l(state, s, code)

    def _m2l(self, state: FMMPreparedState, s, code) -> None:
        """Adds every far pair's M2L (per gear with learned orders) into the locals."""
        topo = state.topology
        # library-wide engine of the float32 M2L rotations, set before every M2L pass
        nat.call("jfm_set_m2l_engine", nat.M2L_ENGINES[self.m2l_engine])
        sched = state.extra.get("gears")
        if sched is not None:
            for i, g in enumerate(sched.gears):
                if sched.flat[i] is not None:
                    launch_flat(sched.flat[i], s, code, state.order, state.multipoles, state.centers4, state.locals)
                elif sched.num_targets[i]:
                    nat.call("jfm_m2l_gear", s, code, state.order, g, 1, sched.targets[i].data_ptr(),
                             sched.num_targets[i], sched.offsets[i].data_ptr(), sched.sources[i].data_ptr(),
                             state.multipoles.data_ptr(), state.centers4.data_ptr(), state.locals.data_ptr())
        elif state.extra.get("flat") is not None:
            launch_flat(state.extra["flat"], s, code, state.order, state.multipoles, state.centers4, state.locals)
        elif topo.m2l_targets.size:
            nat.call("jfm_m2l", s, code, state.order, state.m2l_targets.data_ptr(), int(topo.m2l_targets.size),
                     state.far_off.data_ptr(), state.far_src.data_ptr(), state.multipoles.data_ptr(),
                     state.centers4.data_ptr(), state.locals.data_ptr())

    @staticmethod
    def _l2l(state: FMMPreparedState, s, code) -> None:
        """Pushes the locals down the tree (float64 top levels in mixed precision)."""
        topo = state.topology
        mixed = state.extra.get("mixed")
        first_lev = 0
        if mixed is not None and mixed["ntop"]:
            oh = mixed["l2l_off_host"]
            nat.call("jfm_l2l_top", s, state.order, mixed["l2l_nodes"].data_ptr(), mixed["l2l_off"].data_ptr(),
                     len(oh) - 1, nat.ptr(oh), mixed["top"].data_ptr(), mixed["ntop"], state.parent.data_ptr(),
                     mixed["shadow"].data_ptr(), state.centers4.data_ptr(), mixed["centers64"].data_ptr(),
                     state.locals.data_ptr(), mixed["loc64"].data_ptr())
            first_lev = 3  # depths 1..3 done in float64
        off = topo.l2l_off
        base = state.l2l_nodes.data_ptr()
        for lev in range(first_lev, len(off) - 1):
            a, b = int(off[lev]), int(off[lev + 1])
            if b > a:
                nat.call("jfm_l2l_level", s, code, state.order, base + 4 * a, b - a, state.parent.data_ptr(),
                         state.centers4.data_ptr(), state.locals.data_ptr())

    # ------------------------------------------------------------- evaluate
    def _evaluate_device(self, state: FMMPreparedState, acc: torch.Tensor, pot: Optional[torch.Tensor],
                         flags: Optional[int] = None) -> None:
        topo = state.topology
        if flags is None:
            flags = nat.EVAL_NEAR | nat.EVAL_FAR
            if state.compensated:
                flags |= nat.EVAL_COMPENSATED
        if pot is not None:
            flags |= nat.EVAL_POTENTIAL
        s = nat.current_stream()
        code = nat.dtype_code(state.dtype)
        potp = pot.data_ptr() if pot is not None else None
        items = state.extra.get("p2p_items")
        if items is None:
            nat.call("jfm_evaluate", s, code, state.order, state.pos4.data_ptr(), topo.num_particles,
                     topo.num_leaves, topo.leaf_size, topo.num_internal, state.near_off.data_ptr(),
                     state.near_src.data_ptr(), state.locals.data_ptr(), state.centers4.data_ptr(),
                     state.softening ** 2, state.G, state.perm.data_ptr(), acc.data_ptr(), potp, flags)
            return
        # near field first, far field added onto its rows: the order of
        # step_device, so both round a*G + (-G)*g the same way for any G
        near = bool(flags & nat.EVAL_NEAR)
        if near:
            self._p2p_items(state, items, acc, pot, flags, accumulate=False)
        if flags & nat.EVAL_FAR:
            nat.call("jfm_l2p", s, code, state.order, state.pos4.data_ptr(), topo.num_particles, topo.leaf_size,
                     topo.num_internal, state.locals.data_ptr(), state.centers4.data_ptr(), state.G,
                     state.perm.data_ptr(), acc.data_ptr(), potp, 1 if near else 0)

    @staticmethod
    def _p2p_items(state: FMMPreparedState, items, acc: torch.Tensor, pot: Optional[torch.Tensor], flags: int,
                   accumulate: bool) -> None:
        """P2P over the state's work items, then the partial sums of split leaves."""
        topo = state.topology
        s = nat.current_stream()
        code = nat.dtype_code(state.dtype)
        potp = pot.data_ptr() if pot is not None else None
        nat.call("jfm_p2p_items", s, code, state.pos4.data_ptr(), topo.num_particles, topo.num_leaves,
                 topo.leaf_size, items.num_items, items.leaf.data_ptr(), items.s0.data_ptr(), items.s1.data_ptr(),
                 items.part.data_ptr(), items.num_part, items.partials.data_ptr(), state.near_off.data_ptr(),
                 state.near_src.data_ptr(), state.softening ** 2, state.G, state.perm.data_ptr(), acc.data_ptr(),
                 potp, flags, 1 if accumulate else 0)
        nat.call("jfm_p2p_reduce", s, code, state.pos4.data_ptr(), topo.num_particles, topo.leaf_size,
                 items.num_split, items.split_leaf.data_ptr(), items.split_p0.data_ptr(), items.split_p1.data_ptr(),
                 items.partials.data_ptr(), state.softening ** 2, state.G, state.perm.data_ptr(), acc.data_ptr(),
                 potp, flags, 1 if accumulate else 0)

    def step_device(self, state: FMMPreparedState, positions: torch.Tensor, masses: torch.Tensor,
                    acc: torch.Tensor, pot: Optional[torch.Tensor] = None, events: Optional[dict] = None,
                    overlap: bool = True) -> None:
        """One fixed-topology step (refresh_prepared_state + the evaluation) with the
        near field overlapped: after the gather, P2P runs on a second stream
        while the upward and downward passes run on the current one, then the
        far field is added to the near-field rows.  ``acc``/``pot`` are written
        in input order.  ``events`` (optional): dict of torch.cuda.Event pairs
        'upward', 'downward', 'l2p', 'combine' (current stream) and 'p2p' (its
        stream) recorded around each part; 'l2p' spans the L2P launch only,
        'combine' the join with the P2P stream and the far + near add.
        The zeroing of the locals and the P2M run before the fork, alone on
        the GPU: beside the P2P's workgroups they took 2.5x as long.
        ``overlap=False`` runs the same launches on the current stream only,
        P2P after the downward pass, then an accumulating L2P.  In the
        overlapped order L2P runs before the join into per-state rows
        (unscaled) that jfm_l2p_combine adds to the near-field rows after it,
        with the accumulating L2P's expression (acc += -G * far), so both
        orders, and _evaluate_device (also near first, then far), give the
        same bits for any G."""
        topo = state.topology
        main = torch.cuda.current_stream(state.device)
        s = nat.current_stream()
        code = nat.dtype_code(state.dtype)
        src_code = nat.dtype_code(positions.dtype)
        positions = positions.contiguous()
        masses = masses.to(positions.dtype).contiguous()

        def mark(name, k):
            if events is not None:
                events[name][k].record()

        mark("upward", 0)
        nat.call("jfm_gather_sorted", s, src_code, code, positions.data_ptr(), masses.data_ptr(),
                 state.perm.data_ptr(), topo.num_particles, state.pos4.data_ptr())
        flags = nat.EVAL_COMPENSATED if state.compensated else 0

        def near():
            mark("p2p", 0)
            self._p2p_device(state, acc, pot, flags, accumulate=False)
            mark("p2p", 1)

        early = overlap and self._p2m_before_near
        if overlap:
            # launches that fill the whole GPU run before the fork, alone
            # instead of beside the P2P's workgroups: the zeroing of the
            # locals (and, with _p2m_before_near, the P2M)
            state.locals.zero_()
            if early:
                self._p2m(state, s, code)
            side = self._side_stream(state.device)
            side.wait_stream(main)
            with torch.cuda.stream(side):
                near()
        self._upward(state, s, code, p2m=not early)
        mark("upward", 1)
        mark("downward", 0)
        self._downward(state, s, code, zeroed=overlap)
        mark("downward", 1)
        if overlap:
            # the far field is evaluated while the near field may still run:
            # L2P with G = -1 writes unscaled rows, jfm_l2p_combine adds them
            # after the join exactly as an accumulating L2P would
            far_acc, far_pot = self._far_rows(state)
            mark("l2p", 0)
            nat.call("jfm_l2p", s, code, state.order, state.pos4.data_ptr(), topo.num_particles, topo.leaf_size,
                     topo.num_internal, state.locals.data_ptr(), state.centers4.data_ptr(), -1.0,
                     state.perm.data_ptr(), far_acc.data_ptr(), far_pot.data_ptr() if pot is not None else None, 0)
            mark("l2p", 1)
            mark("combine", 0)
            main.wait_stream(side)
            nat.call("jfm_l2p_combine", s, code, topo.num_particles, state.G, far_acc.data_ptr(),
                     far_pot.data_ptr() if pot is not None else None, acc.data_ptr(),
                     pot.data_ptr() if pot is not None else None)
            mark("combine", 1)
        else:
            near()
            mark("l2p", 0)
            nat.call("jfm_l2p", s, code, state.order, state.pos4.data_ptr(), topo.num_particles, topo.leaf_size,
                     topo.num_internal, state.locals.data_ptr(), state.centers4.data_ptr(), state.G,
                     state.perm.data_ptr(), acc.data_ptr(), pot.data_ptr() if pot is not None else None, 1)
            mark("l2p", 1)
            mark("combine", 0)
            mark("combine", 1)

    @staticmethod
    def _far_rows(state: FMMPreparedState):
        """Per-state rows (accelerations and potentials) for the deferred far
        field of the overlapped step.  Allocated once per state and never
        replaced: a captured EvaluationGraph holds their addresses."""
        rows = state.extra.get("far_rows")
        if rows is None:
            n = state.topology.num_particles
            rows = (torch.empty((n, 3), dtype=state.dtype, device=state.device),
                    torch.empty((n,), dtype=state.dtype, device=state.device))
            state.extra["far_rows"] = rows
        return rows

    @staticmethod
    def _side_stream(device) -> torch.cuda.Stream:
        """The P2P stream of the overlapped step: one per device and process,
        shared by every instance.  torch hands out its pool streams round
        robin and HIP spreads streams over a few hardware queues (4 by
        default), so a fresh stream per instance lands, for some instances
        (the 7th in a row on MI355X), on the hardware queue of the current
        stream, and the P2P then runs after the tree passes instead of beside
        them (8.06 against 7.06-7.11 ms/step on the 4M Plummer workload)."""
        st = _SIDE_STREAMS.get(device)
        if st is None:
            st = _SIDE_STREAMS[device] = torch.cuda.Stream(device=device)  # default priority
        return st

    def _p2p_device(self, state: FMMPreparedState, acc: torch.Tensor, pot: Optional[torch.Tensor], flags: int,
                    accumulate: bool) -> None:
        """Near field only (as jfm_p2p), through the work items when the state has them."""
        items = state.extra.get("p2p_items")
        if pot is not None:
            flags |= nat.EVAL_POTENTIAL
        if items is not None:
            self._p2p_items(state, items, acc, pot, flags, accumulate)
            return
        topo = state.topology
        nat.call("jfm_p2p", nat.current_stream(), nat.dtype_code(state.dtype), state.pos4.data_ptr(),
                 topo.num_particles, topo.num_leaves, topo.leaf_size, state.near_off.data_ptr(),
                 state.near_src.data_ptr(), state.softening ** 2, state.G, state.perm.data_ptr(), acc.data_ptr(),
                 pot.data_ptr() if pot is not None else None, flags, 1 if accumulate else 0)

    def evaluate_prepared_state(self, state: FMMPreparedState, *, target_indices=None,
                                return_potential: bool = False, jit_traversal=None,
                                max_acc_derivative_order: int = 0):
        """Accelerations (and potentials) in input order (solver.py:1292)."""
        if max_acc_derivative_order != 0:
            raise NotImplementedError("acceleration derivatives need the solidfmm basis, which is not built")
        n = state.num_particles
        acc = torch.empty((n, 3), dtype=state.dtype, device=state.device)
        pot = torch.empty((n,), dtype=state.dtype, device=state.device) if return_potential else None
        self._evaluate_device(state, acc, pot)
        if target_indices is not None:
            idx = torch.as_tensor(np.asarray(target_indices) if not isinstance(target_indices, torch.Tensor)
                                  else target_indices, device=state.device).long()
            if idx.ndim != 1:
                raise ValueError("target_indices must be 1-D")
            if idx.numel() and (int(idx.min()) < 0 or int(idx.max()) >= n):
                raise ValueError("target_indices out of range")
            acc = acc.index_select(0, idx)
            if pot is not None:
                pot = pot.index_select(0, idx)
        acc = self._to_output(state, acc)
        if return_potential:
            return acc, self._to_output(state, pot)
        return acc

    def _to_output(self, state: FMMPreparedState, t: torch.Tensor):
        if state.input_kind == "numpy":
            out_dtype = np.dtype(state.input_dtype) if np.issubdtype(np.dtype(state.input_dtype), np.floating) \
                else np.float64
            return t.detach().cpu().numpy().astype(out_dtype, copy=False)
        out_dtype = state.input_dtype if state.input_dtype in (torch.float32, torch.float64) else t.dtype
        return t.to(device=state.input_device, dtype=out_dtype)

    def capture_evaluation(self, state: FMMPreparedState, positions: torch.Tensor, masses: torch.Tensor,
                           return_potential: bool = False) -> "EvaluationGraph":
        """Capture one fixed-topology evaluation (gather, upward, downward, L2P,
        P2P) into a HIP graph.  ``positions``/``masses`` become the graph's
        static input buffers (update them in place between replays); the
        outputs are static tensors in input order."""
        return EvaluationGraph(self, state, positions, masses, return_potential)

    def differentiable_accelerations(self, state: FMMPreparedState, positions, masses, *, target_indices=None,
                                     jit_traversal: bool = False, grad_plan=None, grad_config=None):
        """``(N, 3)`` accelerations (``(T, 3)`` with ``target_indices``) at the
        state's frozen topology, differentiable w.r.t. ``positions`` and
        ``masses`` (solver.py:1358-1368): torch.autograd with libjfm's adjoint
        kernels, expansions recomputed in the backward pass.  All particles
        remain sources.  ``jit_traversal`` and ``grad_config``
        (:class:`jaccpot_amd.config.GradConfig`) select the reference's JAX
        lanes and have no effect here; ``grad_plan`` (large-N only) is not used.
        :meth:`differentiable_accelerations_and_potential` also returns the
        potentials."""
        from .config import resolve_grad_config

        resolve_grad_config(grad_config, grad_plan)
        del jit_traversal
        acc, _ = self._differentiable(state, positions, masses, target_indices, return_potential=False)
        return acc

    def differentiable_accelerations_and_potential(self, state: FMMPreparedState, positions, masses, *,
                                                   target_indices=None, grad_config=None):
        """(acc, pot) of :meth:`differentiable_accelerations`, both differentiable
        (an addition of this build: the reference's method returns accelerations only)."""
        from .config import resolve_grad_config

        resolve_grad_config(grad_config, None)
        return self._differentiable(state, positions, masses, target_indices, return_potential=True)

    def _differentiable(self, state, positions, masses, target_indices, return_potential: bool):
        from .autodiff import differentiable_accelerations

        out = differentiable_accelerations(self, state, positions, masses, return_potential=True)
        acc, pot = out
        if target_indices is not None:
            idx = torch.as_tensor(np.asarray(target_indices) if not isinstance(target_indices, torch.Tensor)
                                  else target_indices, device=acc.device).long()
            if idx.ndim != 1:
                raise ValueError("target_indices must be 1-D")
            if idx.numel() and (int(idx.min()) < 0 or int(idx.max()) >= acc.shape[0]):
                raise ValueError("target_indices out of range")
            acc = acc.index_select(0, idx)
            pot = pot.index_select(0, idx)
        return acc, (pot if return_potential else None)

    def compute_accelerations(self, positions, masses, *, target_indices=None, bounds=None, leaf_size: int = 16,
                              max_order: int = 4, return_potential: bool = False, theta: Optional[float] = None,
                              reuse_prepared_state: bool = False, max_acc_derivative_order: int = 0):
        """prepare_state + evaluate_prepared_state (solver.py:917).

        ``reuse_prepared_state``: evaluate the state this method prepared last
        again when the position and mass objects and the preparation arguments
        are the same objects / values as in that call (an identity check, like
        the reference's: an array mutated in place is not noticed).  The state
        is cached with its key (fmm_derivatives.py:141-185), so direct
        ``prepare_state`` calls in between do not change what is reused."""
        args = (int(leaf_size), int(max_order), theta)
        last = self._last_inputs
        if (reuse_prepared_state and bounds is None and last is not None
                and last[0]() is positions and last[1]() is masses and last[2] == args):
            state = last[3]
        else:
            state = self.prepare_state(positions, masses, bounds=bounds, leaf_size=leaf_size, max_order=max_order,
                                       theta=theta)
            try:
                self._last_inputs = (weakref.ref(positions), weakref.ref(masses), args, state)
            except TypeError:  # objects without weak references (lists) are never reused
                self._last_inputs = None
        return self.evaluate_prepared_state(state, target_indices=target_indices, return_potential=return_potential,
                                            max_acc_derivative_order=max_acc_derivative_order)


class EvaluationGraph:
    """A captured fixed-topology evaluation; ``replay()`` re-runs it on the
    current contents of ``positions``/``masses`` and returns (acc[, pot])."""

    def __init__(self, fmm: FastMultipoleMethod, state: FMMPreparedState, positions: torch.Tensor,
                 masses: torch.Tensor, return_potential: bool = False):
        if positions.device != state.device or masses.device != state.device:
            raise ValueError("graph inputs must live on the state's device")
        self.positions = positions
        self.masses = masses
        n = state.num_particles
        self.acc = torch.empty((n, 3), dtype=state.dtype, device=state.device)
        self.pot = torch.empty((n,), dtype=state.dtype, device=state.device) if return_potential else None

        def step():
            fmm.step_device(state, self.positions, self.masses, self.acc, self.pot)

        side = torch.cuda.Stream(device=state.device)
        side.wait_stream(torch.cuda.current_stream(state.device))
        with torch.cuda.stream(side):
            step()  # eager run on the capture stream first (allocations settle)
        torch.cuda.current_stream(state.device).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            step()
        # the replay reads and writes these buffers by address: keep them alive
        self.state = state
        self.far_rows = state.extra.get("far_rows")

    def replay(self):
        self.graph.replay()
        return (self.acc, self.pot) if self.pot is not None else self.acc
