"""The reference's configuration surface, resolved for this path.

``FastMultipoleMethod`` accepts every keyword of the reference's constructor
(jaccpot/solver.py:664-692, plus the legacy keywords its constructor pops,
solver.py:693-905), of ``prepare_state`` (solver.py:1012-1024) and of
``differentiable_accelerations`` (solver.py:1358-1368), and the config objects
``GradConfig`` (jaccpot/config.py:420-516) and ``FMMAdvancedConfig``
(config.py:518-554).  Each option falls in one of three groups:

* execution and memory knobs (kernel selection, chunk sizes, caches, JIT
  switches, queue sizes, schedules): they do not change the arithmetic of the
  fixed-topology real-basis path, so they are accepted, validated where the
  reference validates them, and otherwise have no effect -- every step here
  runs in libjfm's HIP kernels;
* options that would select a different algorithm (another tree, a grouped
  M2L, a different acceptance model or node geometry): accepted at the values
  this path implements, ``NotImplementedError`` otherwise;
* options outside their documented domain: ``ValueError``, as in the
  reference; unknown keywords: ``TypeError``.

The adaptive-order options of the reference (``mac_force_scale_*``,
``adaptive_error_model``, ``adaptive_eps``, ``dehnen_geometry_mode``,
``mac_theta_max``) tune the hand-written error model behind its per-pair gear
tags (runtime/_adaptive_policy.py:2216-2384).  Here a learned order model
chooses the gears (``jaccpot_amd.learned``) and the plain fixed-order FMM at
the top gear judges it on sampled targets; with ``adaptive_order=True`` only
their default values are accepted, without it they are inert as in the
reference.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Optional

# --------------------------------------------------------------- config objects


@dataclass(frozen=True)
class GradConfig:
    """Execution options of ``differentiable_accelerations`` (jaccpot/config.py:420-516).

    Field names, defaults and the validation of ``nearfield_lane`` follow the
    reference.  Every field selects how the reference's JAX program traverses
    or transposes the same sums (lane, tiling, analytic or autodiff reverse
    rule); the adjoint kernels here have one schedule, so the fields are
    recorded and do not change the result."""

    nearfield_lane: str = "auto"
    nearfield_fast_lane_min_particles: int = 100_000
    fused_m2l_pallas: Optional[bool] = None
    analytic_p2p_vjp: Optional[bool] = None
    analytic_l2p_vjp: Optional[bool] = None
    reverse_tiers: Optional[int] = None
    reverse_tier_min_gain: Optional[float] = None
    reverse_skip_empty_tiles: Optional[bool] = None
    reverse_leaf_batch: Optional[int] = None
    reverse_block_tile: Optional[int] = None

    def __post_init__(self) -> None:
        lane = str(self.nearfield_lane).strip().lower()
        if lane not in ("auto", "bucketed", "fast_lane"):
            raise ValueError(f"GradConfig.nearfield_lane must be 'auto', 'bucketed' or 'fast_lane'; "
                             f"got {self.nearfield_lane!r}")
        if int(self.nearfield_fast_lane_min_particles) < 0:
            raise ValueError("GradConfig.nearfield_fast_lane_min_particles must be >= 0; got "
                             f"{self.nearfield_fast_lane_min_particles!r}")


@dataclass(frozen=True)
class TreeConfig:
    """Tree-construction overrides (jaccpot/config.py:107-139)."""

    tree_type: Optional[str] = None
    mode: Optional[str] = None
    leaf_target: Optional[int] = None
    refine_local: Optional[bool] = None
    max_refine_levels: Optional[int] = None
    aspect_threshold: Optional[float] = None


@dataclass(frozen=True)
class FarFieldConfig:
    """Far-field overrides (jaccpot/config.py:140-209)."""

    grouped_interactions: Optional[bool] = None
    mode: str = "auto"
    rotation: Optional[str] = None
    m2l_chunk_size: Optional[int] = None
    l2l_chunk_size: Optional[int] = None
    streamed_far_pairs: Optional[bool] = None
    mixed_order: bool = False
    mixed_order_min_order: Optional[int] = None
    retain_far_pairs_for_grad: bool = False


@dataclass(frozen=True)
class NearFieldConfig:
    """Near-field overrides (jaccpot/config.py:210-248)."""

    mode: str = "auto"
    edge_chunk_size: int = 256
    precompute_scatter_schedules: bool = True


@dataclass(frozen=True)
class RuntimePolicyConfig:
    """Execution-policy overrides (jaccpot/config.py:330-418); none changes the arithmetic."""

    execution_backend: str = "auto"
    host_refine_mode: str = "auto"
    fail_fast: bool = False
    jit_tree: Optional[bool] = None
    jit_traversal: Optional[bool] = None
    memory_objective: str = "balanced"
    memory_budget_bytes: Optional[int] = None
    max_pair_queue: Optional[int] = None
    pair_process_block: Optional[int] = None
    traversal_config: Optional[Any] = None
    enable_interaction_cache: bool = True
    retain_traversal_result: bool = True
    retain_interactions: bool = True
    prepare_stage_memory_split_enabled: Optional[bool] = None
    autotune_m2l_chunk: bool = False
    precompute_grouped_class_segments: Optional[bool] = None
    grouped_schedule_budget_bytes: Optional[int] = None
    nearfield_schedule_item_cap: Optional[int] = None
    upward_leaf_batch_size: Optional[int] = None


@dataclass(frozen=True)
class FMMAdvancedConfig:
    """Aggregate of the override groups (jaccpot/config.py:518-554)."""

    tree: TreeConfig = TreeConfig()
    farfield: FarFieldConfig = FarFieldConfig()
    nearfield: NearFieldConfig = NearFieldConfig()
    runtime: RuntimePolicyConfig = RuntimePolicyConfig()
    mac_type: Optional[str] = None
    dehnen_radius_scale: float = 1.0


# ------------------------------------------------------------ option resolution

# Legacy constructor keywords the reference pops (solver.py:693-905) that only
# select execution strategy or memory policy.
INERT_LEGACY = frozenset({
    "runtime_path", "execution_backend", "tree_build_mode", "max_refine_levels", "aspect_threshold",
    "retain_far_pairs_for_grad", "streamed_far_pairs", "mixed_order_min_order", "m2l_chunk_size",
    "l2l_chunk_size", "farfield_mode", "nearfield_mode", "nearfield_edge_chunk_size",
    "precompute_nearfield_scatter_schedules", "host_refine_mode", "fail_fast", "memory_objective",
    "memory_budget_bytes", "max_pair_queue", "pair_process_block", "use_dense_interactions", "traversal_config",
    "enable_interaction_cache", "retain_traversal_result", "retain_interactions",
    "prepare_stage_memory_split_enabled", "autotune_m2l_chunk", "precompute_grouped_class_segments",
    "grouped_schedule_budget_bytes", "nearfield_schedule_item_cap", "upward_leaf_batch_size", "complex_rotation",
    "fixed_max_leaf_size",
})

_FORCE_SCALE_MODES = ("prev", "prepass", "paper", "paper_cached", "paper_fb", "paper_fb_cached")
_ERROR_MODELS = ("tail_proxy", "dehnen_degree", "dehnen_paper")
_GEOMETRY_MODES = ("com", "exact", "tree", "tree_approx", "runtime")


def _algorithm_checks(tree_type, refine_local, leaf_target, grouped, mixed_order, mac_type, radius_scale,
                      fixed_order, rotation) -> None:
    if tree_type not in (None, "radix"):
        raise NotImplementedError(f"tree_type={tree_type!r}: only the static radix tree is implemented")
    if refine_local:
        raise NotImplementedError("refine_local=True (host leaf refinement) is not implemented; "
                                  "the static radix tree is used")
    if leaf_target is not None:
        raise NotImplementedError("target_leaf_particles / leaf_target is not implemented; pass leaf_size")
    if grouped:
        raise NotImplementedError("grouped_interactions=True (one rotation per displacement class) is not "
                                  "implemented: every far pair is translated by its own displacement")
    if mixed_order:
        raise NotImplementedError("mixed_order far field is not implemented; use adaptive_order=True with p_gears")
    if mac_type is not None and str(mac_type).strip().lower() != "dehnen":
        raise NotImplementedError(f"mac_type={mac_type!r}: only the Dehnen sum-form MAC "
                                  "(r_t + r_s)^2 <= theta^2 d^2 is implemented")
    if float(radius_scale) != 1.0:
        raise NotImplementedError("dehnen_radius_scale != 1 is not implemented")
    if fixed_order is not None:
        raise NotImplementedError("fixed_order overrides are not implemented; pass max_order")
    if rotation not in (None, "solidfmm", "cached"):
        raise NotImplementedError(f"far-field rotation={rotation!r} is not implemented")


def resolve_constructor_options(*, adaptive_order: bool, use_pallas, mac_force_scale_mode: str,
                                mac_force_scale_prepass_theta, mac_force_scale_fb_inflation: float,
                                adaptive_error_model: str, adaptive_eps, dehnen_geometry_mode: str,
                                mac_theta_max: float, advanced, legacy_kwargs: dict) -> dict:
    """Validate the reference-only constructor options; returns the normalised
    values that the facade records.  Raises ValueError / NotImplementedError /
    TypeError as described in the module docstring."""
    legacy = dict(legacy_kwargs)
    if use_pallas is not None and not isinstance(use_pallas, bool):
        raise ValueError("use_pallas must be None, True or False")
    mode = str(mac_force_scale_mode).strip().lower()
    if mode not in _FORCE_SCALE_MODES:
        raise ValueError("mac_force_scale_mode must be 'prev', 'prepass', 'paper', 'paper_cached', 'paper_fb', "
                         "or 'paper_fb_cached'")
    if mac_force_scale_prepass_theta is not None and not (0.0 < float(mac_force_scale_prepass_theta) <= 1.0):
        raise ValueError("mac_force_scale_prepass_theta must be in (0, 1]")
    if float(mac_force_scale_fb_inflation) < 0.0:
        raise ValueError("mac_force_scale_fb_inflation must be >= 0")
    err_model = str(adaptive_error_model).strip().lower()
    if err_model not in _ERROR_MODELS:
        raise ValueError("adaptive_error_model must be 'tail_proxy', 'dehnen_degree', or 'dehnen_paper'")
    geom = str(dehnen_geometry_mode).strip().lower()
    if geom not in _GEOMETRY_MODES:
        raise ValueError("dehnen_geometry_mode must be 'com', 'exact', 'tree', 'tree_approx', or 'runtime'")
    if not (0.0 < float(mac_theta_max) <= 1.0):
        raise ValueError("mac_theta_max must be in (0, 1]")
    if adaptive_eps is not None and float(adaptive_eps) <= 0.0:
        raise ValueError("adaptive_eps must be > 0 when provided")
    if adaptive_order:
        changed = [name for name, val, default in (
            ("mac_force_scale_mode", mode, "prev"), ("mac_force_scale_prepass_theta", mac_force_scale_prepass_theta,
                                                     None),
            ("mac_force_scale_fb_inflation", float(mac_force_scale_fb_inflation), 1.0),
            ("adaptive_error_model", err_model, "tail_proxy"), ("adaptive_eps", adaptive_eps, None),
            ("dehnen_geometry_mode", geom, "com"), ("mac_theta_max", float(mac_theta_max), 1.0)) if val != default]
        if changed:
            raise NotImplementedError(
                f"{', '.join(changed)}: these tune the reference's hand-written adaptive acceptance model; here "
                "the per-pair gears come from the learned order model, judged against the plain top-gear FMM "
                "(leave them at their defaults)")

    adv = advanced
    if adv is not None:
        for attr in ("tree", "farfield", "nearfield", "runtime"):
            if not hasattr(adv, attr):
                raise TypeError(f"advanced must be an FMMAdvancedConfig (missing .{attr})")
        tree, far = adv.tree, adv.farfield
        _algorithm_checks(getattr(tree, "tree_type", None), getattr(tree, "refine_local", None),
                          getattr(tree, "leaf_target", None), getattr(far, "grouped_interactions", None),
                          getattr(far, "mixed_order", False), getattr(adv, "mac_type", None),
                          getattr(adv, "dehnen_radius_scale", 1.0), None, getattr(far, "rotation", None))

    # legacy keywords: the common expanse-style names map onto constructor
    # arguments the facade already has; the rest are checked or inert
    renamed = {}
    if "expansion_basis" in legacy:
        renamed["basis"] = legacy.pop("expansion_basis")
    if "preset" in legacy:
        renamed["preset"] = legacy.pop("preset")
    _algorithm_checks(legacy.pop("tree_type", None), legacy.pop("refine_local", None),
                      legacy.pop("target_leaf_particles", None), legacy.pop("grouped_interactions", None),
                      legacy.pop("mixed_order_farfield", False), legacy.pop("mac_type", None),
                      legacy.pop("dehnen_radius_scale", 1.0), legacy.pop("fixed_order", None), None)
    inert = sorted(k for k in legacy if k in INERT_LEGACY)
    for k in inert:
        legacy.pop(k)
    if legacy:
        raise TypeError(f"Unknown jaccpot.FastMultipoleMethod kwargs: {', '.join(sorted(legacy))}")
    return {"use_pallas": use_pallas, "mac_force_scale_mode": mode, "adaptive_error_model": err_model,
            "adaptive_eps": adaptive_eps, "dehnen_geometry_mode": geom, "mac_theta_max": float(mac_theta_max),
            "advanced": adv, "inert_legacy": inert, "renamed": renamed}


def resolve_prepare_options(*, adaptive_order: bool, cache_policy: str, force_scale_nodes,
                            runtime_overrides_override, fused_device_mode: bool, extra: dict) -> None:
    """prepare_state's reference-only keywords (solver.py:1012-1024): cache
    policy, runtime overrides and the fused device layout are execution
    choices; a per-node force scale only feeds the reference's adaptive
    acceptance model, which this path replaces."""
    extra = dict(extra)
    extra.pop("retain_far_pairs_for_grad", None)  # the far list is always kept on the state
    if extra:
        raise TypeError(f"prepare_state() got unexpected keyword arguments: {', '.join(sorted(extra))}")
    if not isinstance(cache_policy, str):
        raise ValueError("cache_policy must be a string")
    if force_scale_nodes is not None and adaptive_order:
        raise NotImplementedError("force_scale_nodes feeds the reference's adaptive acceptance model; the learned "
                                  "order model here does not take a force scale")
    del runtime_overrides_override, fused_device_mode


def resolve_grad_config(grad_config, grad_plan) -> Optional[GradConfig]:
    """Validate a GradConfig (or any object with its fields); grad_plan is the
    reference's large-N plan object and has no counterpart here."""
    if grad_plan is not None:
        raise NotImplementedError("grad_plan (a LargeNGradPlan of the reference's large-N lane) is not used: "
                                  "the adjoint kernels here need no plan")
    if grad_config is None:
        return None
    if isinstance(grad_config, GradConfig):
        return grad_config
    fields = GradConfig.__dataclass_fields__
    return GradConfig(**{k: getattr(grad_config, k) for k in fields if hasattr(grad_config, k)})
