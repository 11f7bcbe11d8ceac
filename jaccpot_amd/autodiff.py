"""Differentiable FMM at a frozen topology (reverse mode on libjfm kernels).

Counterpart of ``differentiable_accelerations`` (jaccpot/runtime/fmm_evaluate.py:733,
:542-731): the upward and downward passes are re-run on the live positions and
masses with the prepared state's frozen topology, and gradients with respect
to positions and masses come from hand-written adjoint kernels (P2P, L2P, M2L
and the tree passes), recomputing every expansion in the backward pass
instead of storing it.  ``torch.autograd.Function`` plays the role the
reference gives ``jax.custom_vjp``.
"""

from __future__ import annotations

import torch

from . import _native as nat


def _check_state(state):
    if state.extra.get("gears") is not None or state.extra.get("mixed") is not None:
        raise NotImplementedError("gradients are implemented for the fixed-order, single-precision-mode path")
    if state.order > 7:
        raise NotImplementedError("adjoint kernels exist for orders 1..7")


def _recompute(fmm, state, positions: torch.Tensor, masses: torch.Tensor):
    """Forward expansions at the live inputs; returns (L_own, L_full)."""
    topo = state.topology
    s = nat.current_stream()
    code = nat.dtype_code(state.dtype)
    nat.call("jfm_gather_sorted", s, nat.dtype_code(positions.dtype), code, positions.data_ptr(),
             masses.data_ptr(), state.perm.data_ptr(), topo.num_particles, state.pos4.data_ptr())
    fmm._upward(state, s, code)
    state.locals.zero_()
    if topo.m2l_targets.size:
        nat.call("jfm_m2l", s, code, state.order, state.m2l_targets.data_ptr(), int(topo.m2l_targets.size),
                 state.far_off.data_ptr(), state.far_src.data_ptr(), state.multipoles.data_ptr(),
                 state.centers4.data_ptr(), state.locals.data_ptr())
    lown = state.locals.clone()
    off = topo.l2l_off
    base = state.l2l_nodes.data_ptr()
    for lev in range(len(off) - 1):
        a, b = int(off[lev]), int(off[lev + 1])
        if b > a:
            nat.call("jfm_l2l_level", s, code, state.order, base + 4 * a, b - a, state.parent.data_ptr(),
                     state.centers4.data_ptr(), state.locals.data_ptr())
    return lown


def fmm_vjp(fmm, state, positions: torch.Tensor, masses: torch.Tensor, acc_bar: torch.Tensor,
            pot_bar: torch.Tensor):
    """(positions_bar [N,3], masses_bar [N]) for cotangents of (acc [N,3], pot [N]), input order."""
    _check_state(state)
    topo = state.topology
    dt = state.dtype
    dev = state.device
    code = nat.dtype_code(dt)
    s = nat.current_stream()
    x = positions.to(dev, dt).contiguous()
    m = masses.to(dev, dt).contiguous()
    lown = _recompute(fmm, state, x, m)
    n = topo.num_particles
    cot4 = torch.empty((n, 4), dtype=dt, device=dev)
    ab = acc_bar.to(dev, dt).contiguous()
    pb = pot_bar.to(dev, dt).contiguous()
    nat.call("jfm_gather_sorted", s, code, code, ab.data_ptr(), pb.data_ptr(), state.perm.data_ptr(), n,
             cot4.data_ptr())
    grad4 = torch.empty((n, 4), dtype=dt, device=dev)
    items = state.extra.get("p2p_items")
    if items is None:
        nat.call("jfm_p2p_adjoint", s, code, state.pos4.data_ptr(), cot4.data_ptr(), n, topo.num_leaves,
                 topo.leaf_size, state.near_off.data_ptr(), state.near_src.data_ptr(), state.softening ** 2, state.G,
                 grad4.data_ptr())
    else:  # long near lists cut into work items, as in the forward P2P
        nat.call("jfm_p2p_adjoint_items", s, code, state.pos4.data_ptr(), cot4.data_ptr(), n, topo.num_leaves,
                 topo.leaf_size, items.num_items, items.leaf.data_ptr(), items.s0.data_ptr(), items.s1.data_ptr(),
                 items.part.data_ptr(), items.num_part, items.partials.data_ptr(), state.near_off.data_ptr(),
                 state.near_src.data_ptr(), state.softening ** 2, state.G, grad4.data_ptr())
        nat.call("jfm_p2p_adjoint_reduce", s, code, cot4.data_ptr(), n, topo.leaf_size, items.num_split,
                 items.split_leaf.data_ptr(), items.split_p0.data_ptr(), items.split_p1.data_ptr(),
                 items.partials.data_ptr(), state.softening ** 2, state.G, grad4.data_ptr())
    lbar = torch.zeros_like(state.locals)
    nat.call("jfm_leaf_adjoint_a", s, code, state.order, state.pos4.data_ptr(), cot4.data_ptr(), n, topo.leaf_size,
             topo.num_internal, state.locals.data_ptr(), state.centers4.data_ptr(), state.G, grad4.data_ptr(),
             lbar.data_ptr())
    # cotangents of the locals up the tree (l2l^T = m2m on parity-flipped rows)
    off = topo.m2m_off
    for lev in range(len(off) - 1):
        a, b = int(off[lev]), int(off[lev + 1])
        if b > a:
            nat.call("jfm_m2m_level", s, code, state.order, state.m2m_nodes.data_ptr() + 4 * a, b - a,
                     state.left.data_ptr(), state.right.data_ptr(), state.node_range.data_ptr(), lbar.data_ptr(),
                     state.centers4.data_ptr())
    mbar = torch.zeros_like(state.multipoles)
    cbar4 = torch.zeros((topo.num_nodes, 4), dtype=dt, device=dev)
    if topo.m2l_targets.size:
        nt = int(topo.m2l_targets.size)
        nat.call("jfm_m2l", s, code, state.order, state.m2l_targets.data_ptr(), nt, state.far_off.data_ptr(),
                 state.far_src.data_ptr(), lbar.data_ptr(), state.centers4.data_ptr(), mbar.data_ptr())
        nat.call("jfm_center_adjoint", s, code, state.order, state.m2l_targets.data_ptr(), nt,
                 state.far_off.data_ptr(), state.far_src.data_ptr(), state.multipoles.data_ptr(), lown.data_ptr(),
                 lbar.data_ptr(), mbar.data_ptr(), state.centers4.data_ptr(), cbar4.data_ptr())
    # cotangents of the multipoles down the tree (m2m^T = l2l on parity-flipped rows)
    off = topo.l2l_off
    for lev in range(len(off) - 1):
        a, b = int(off[lev]), int(off[lev + 1])
        if b > a:
            nat.call("jfm_l2l_level", s, code, state.order, state.l2l_nodes.data_ptr() + 4 * a, b - a,
                     state.parent.data_ptr(), state.centers4.data_ptr(), mbar.data_ptr())
    chain = torch.empty((topo.num_leaves, 8), dtype=dt, device=dev)
    nat.call("jfm_com_chain", s, code, topo.num_leaves, topo.num_internal, state.parent.data_ptr(),
             state.node_range.data_ptr(), state.centers4.data_ptr(), cbar4.data_ptr(), chain.data_ptr())
    gpos = torch.empty((n, 3), dtype=dt, device=dev)
    gmass = torch.empty((n,), dtype=dt, device=dev)
    nat.call("jfm_leaf_adjoint_b", s, code, state.order, state.pos4.data_ptr(), n, topo.leaf_size, topo.num_internal,
             mbar.data_ptr(), state.centers4.data_ptr(), chain.data_ptr(), grad4.data_ptr(), state.perm.data_ptr(),
             gpos.data_ptr(), gmass.data_ptr())
    return gpos, gmass


class _FMMForce(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, masses, fmm, state):
        x = positions.detach().to(state.device, state.dtype).contiguous()
        m = masses.detach().to(state.device, state.dtype).contiguous()
        fmm.refresh_prepared_state(state, x, m)
        n = state.num_particles
        acc = torch.empty((n, 3), dtype=state.dtype, device=state.device)
        pot = torch.empty((n,), dtype=state.dtype, device=state.device)
        fmm._evaluate_device(state, acc, pot)
        ctx.fmm = fmm
        ctx.state = state
        ctx.save_for_backward(x, m)
        return acc.to(positions.dtype), pot.to(positions.dtype)

    @staticmethod
    def backward(ctx, gacc, gpot):
        x, m = ctx.saved_tensors
        n = x.shape[0]
        if gacc is None:
            gacc = torch.zeros((n, 3), dtype=x.dtype, device=x.device)
        if gpot is None:
            gpot = torch.zeros((n,), dtype=x.dtype, device=x.device)
        gx, gm = fmm_vjp(ctx.fmm, ctx.state, x, m, gacc, gpot)
        return gx, gm, None, None


def differentiable_accelerations(fmm, state, positions: torch.Tensor, masses: torch.Tensor,
                                 return_potential: bool = False):
    """Accelerations (and potentials) in input order with autograd support."""
    acc, pot = _FMMForce.apply(positions, masses, fmm, state)
    return (acc, pot) if return_potential else acc
