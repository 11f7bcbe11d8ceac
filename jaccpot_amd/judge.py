"""Run-time judge of the learned per-pair M2L orders, with the plain
fixed-order FMM as fallback.

The reference accepts a far pair at the highest gear (its safety check) and
then tags it with the smallest passing gear (README.md:290-299,
runtime/_adaptive_policy.py:2216-2384).  Here a learned model picks the gear,
so nothing in the model itself guarantees the result: after the gears are
chosen in ``prepare_state``, the judge evaluates, on a seeded sample of
targets,

* the FMM with the learned per-pair gears, and
* the plain FMM with every far pair at the top gear (same topology, same
  near field, same precision policy),

both against the float64 direct sum at those targets.  When the learned
median relative force error exceeds ``factor`` times the plain one (plus a
float32 noise floor), the state falls back to the plain fixed-order far field:
the gear schedule is dropped and the downward pass runs every pair at the top
gear.  The decision and both error figures are recorded in
``state.extra["order_judge"]``.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from .direct import direct_sum_gravitational_acceleration

# relative force errors below this are float32 rounding, not truncation
NOISE_FLOOR = 2e-6


def _errors(acc: torch.Tensor, ref: torch.Tensor) -> np.ndarray:
    a = acc.to(torch.float64)
    return (torch.linalg.norm(a - ref, dim=1) / torch.linalg.norm(ref, dim=1).clamp_min(1e-300)).cpu().numpy()


def judge_learned_orders(fmm, state, positions: torch.Tensor, masses: torch.Tensor, *, num_targets: int,
                         factor: float, seed: int) -> dict:
    """Judge the learned gears of ``state`` (prepared, gears set, expansions of
    the learned downward pass in place) on ``num_targets`` seeded targets and
    fall back to the top gear for every pair when they lose.  ``positions`` /
    ``masses`` are the device inputs in input order."""
    sched = state.extra.get("gears")
    if sched is None:
        raise ValueError("the state has no learned gears to judge")
    n = state.num_particles
    k = int(min(max(num_targets, 1), n))
    tgt = np.sort(np.random.default_rng(seed).choice(n, size=k, replace=False)).astype(np.int64)
    tgt_d = torch.as_tensor(tgt, device=state.device)
    ref = direct_sum_gravitational_acceleration(positions, masses, G=state.G, softening=state.softening,
                                                target_indices=tgt)

    acc = torch.empty((n, 3), dtype=state.dtype, device=state.device)
    fmm._evaluate_device(state, acc, None)
    err_learned = _errors(acc[tgt_d], ref)
    locals_learned = state.locals.clone()

    # the judge: every far pair at the top gear on the same expansions
    state.extra["gears"] = None
    s = nat.current_stream()
    fmm._downward(state, s, nat.dtype_code(state.dtype))
    fmm._evaluate_device(state, acc, None)
    err_plain = _errors(acc[tgt_d], ref)

    med_l, med_p = float(np.median(err_learned)), float(np.median(err_plain))
    accepted = med_l <= factor * med_p + NOISE_FLOOR
    from .perf import m2l_flops

    verdict = {
        "accepted": bool(accepted), "targets": k, "seed": int(seed), "factor": float(factor),
        "median_learned": med_l, "median_plain": med_p,
        "p99_learned": float(np.percentile(err_learned, 99)), "p99_plain": float(np.percentile(err_plain, 99)),
        "m2l_gflop_learned": sched.m2l_flops / 1e9,
        "m2l_gflop_plain": m2l_flops(state.order) * state.topology.num_far_pairs / 1e9,
    }
    if accepted:
        state.extra["gears"] = sched
        state.locals.copy_(locals_learned)
    else:
        # fallback: the plain fixed-order far field just computed stays
        state.extra.pop("gear_of_pair", None)
        state.extra.pop("gears", None)
    state.extra["order_judge"] = verdict
    return verdict
