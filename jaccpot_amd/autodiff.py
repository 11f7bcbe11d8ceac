"""Differentiable FMM at a frozen topology (reverse mode on libjfm kernels).

Counterpart of ``differentiable_accelerations`` (jaccpot/runtime/fmm_evaluate.py:733,
:542-731): the upward and downward passes are re-run on the live positions and
masses with the prepared state's frozen topology, and gradients with respect
to positions and masses come from hand-written adjoint kernels (P2P, L2P, M2L
and the tree passes), recomputing every expansion in the backward pass
instead of storing it.  ``torch.autograd.Function`` plays the role the
reference gives ``jax.custom_vjp``.

Learned gears: every far pair's M2L and centre terms are differentiated at
the pair's own gear, with the gears held fixed (the network's choice is
piecewise constant in the inputs, as the reference's traced gear tags are).
Mixed precision: the recomputed expansions use the forward's float64 top
levels; the transposed tree passes run in the working dtype.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat


def _check_state(state):
    if not 1 <= state.order <= 8:
        raise NotImplementedError("adjoint kernels exist for orders 1..8")


@dataclass
class AdjointGears:
    """Per-pair orders of a learned-gear state, arranged for the reverse pass.

    ``fwd[k]`` is the gear of far pair k = (u <- v) in CSR order, ``rev[k]``
    the gear of its mirror (v <- u).  ``targets/num_targets/offsets/sources``
    are, per gear g, the transposed CSR of the gear-g pairs: row v lists every
    u whose pair (u <- v) runs at g, so that Mbar_own_v = sum over those u of
    m2l_g(Lbar_u) (m2l_g is symmetric in the packed real basis)."""

    source: object
    gears: tuple
    fwd: torch.Tensor
    rev: torch.Tensor
    targets: list
    num_targets: list
    offsets: list
    sources: list


def adjoint_gears(state) -> AdjointGears:
    """Builds (once per gear assignment) the reverse-pass gear arrays on the device.

    Two radix sorts with the pair index as payload -- of the (u, v) keys and of
    the mirrored (v, u) keys -- put every pair next to its mirror: the far
    lists are symmetric, so both sorted key sequences are the same and
    position t pairs forward pair pf[t] with mirrored pair pr[t]."""
    from .dtopology import _csr, _key_bits, _sort_u64
    from .flatlist import pair_targets

    gop = state.extra["gear_of_pair"]
    cached = state.extra.get("adjoint_gears")
    if cached is not None and cached.source is gop:
        return cached
    topo = state.topology
    dev = state.device
    s = nat.current_stream()
    K = topo.num_far_pairs
    nodes = topo.num_nodes
    gears = tuple(state.extra["gears"].gears)
    fwd = torch.as_tensor(np.asarray(gop)).to(device=dev, dtype=torch.int8).contiguous()
    if fwd.numel() != K:
        raise ValueError(f"gear_of_pair has {fwd.numel()} entries for {K} far pairs")
    rev = torch.empty_like(fwd)
    targets, num_targets, offsets, sources = [], [], [], []
    one = torch.zeros(1, dtype=torch.int32, device=dev)
    if K:
        pt = pair_targets(state)
        ps = state.far_src[:K]
        kf = torch.empty(K, dtype=torch.int64, device=dev)
        kr = torch.empty(K, dtype=torch.int64, device=dev)
        nat.call("jfm_pair_keys", s, K, pt.data_ptr(), ps.data_ptr(), kf.data_ptr())
        nat.call("jfm_pair_keys", s, K, ps.data_ptr(), pt.data_ptr(), kr.data_ptr())
        idx = torch.arange(K, dtype=torch.int32, device=dev)
        end = 32 + _key_bits(max(nodes - 1, 1))
        sf, pf = _sort_u64(kf, end, idx)
        sr, pr = _sort_u64(kr, end, idx)
        del kf, kr, idx
        if not torch.equal(sf, sr):
            raise ValueError("the far pair lists are not symmetric; the M2L adjoint needs every mirror pair")
        del sf
        rev[pr.long()] = fwd[pf.long()]
        gear_t = fwd[pr.long()]  # gear of the forward pair behind each mirrored key
        del pf, pr
    for g in gears:
        if not K:
            targets.append(one), num_targets.append(0), offsets.append(None), sources.append(one)
            continue
        kg = sr[gear_t == g].contiguous()
        off, src = _csr(kg, nodes)
        rows = torch.empty(max(nodes, 1), dtype=torch.int64, device=dev)
        cnt = torch.zeros(1, dtype=torch.int64, device=dev)
        nat.call("jfm_nonempty_rows", s, nodes, off.data_ptr(), rows.data_ptr(), cnt.data_ptr())
        nt = int(cnt.item())
        tl, _ = _sort_u64(rows[:nt], 32)
        targets.append(tl.to(torch.int32) if nt else one)
        num_targets.append(nt)
        offsets.append(off)
        sources.append(src if kg.numel() else one)
    out = AdjointGears(gop, gears, fwd, rev, targets, num_targets, offsets, sources)
    state.extra["adjoint_gears"] = out
    return out


def _recompute(fmm, state, s, code):
    """Forward expansions at the live inputs (already gathered into
    ``state.pos4``) with the forward's own passes (learned gears, float64 top
    levels); returns the locals after M2L only (the full locals are left in
    ``state.locals``)."""
    fmm._upward(state, s, code)
    state.locals.zero_()
    fmm._m2l(state, s, code)
    lown = state.locals.clone()
    fmm._l2l(state, s, code)
    return lown


def _p2p_adjoint(state, code, s, cot4, grad4):
    """Near-field part of the particle gradient into grad4 (written, not added)."""
    topo = state.topology
    n = topo.num_particles
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
    ag = adjoint_gears(state) if state.extra.get("gears") is not None else None
    n = topo.num_particles
    # tree-order particles and cotangents: all that the adjoint P2P reads
    nat.call("jfm_gather_sorted", s, code, code, x.data_ptr(), m.data_ptr(), state.perm.data_ptr(), n,
             state.pos4.data_ptr())
    cot4 = torch.empty((n, 4), dtype=dt, device=dev)
    ab = acc_bar.to(dev, dt).contiguous()
    pb = pot_bar.to(dev, dt).contiguous()
    nat.call("jfm_gather_sorted", s, code, code, ab.data_ptr(), pb.data_ptr(), state.perm.data_ptr(), n,
             cot4.data_ptr())
    grad4 = torch.empty((n, 4), dtype=dt, device=dev)
    # the adjoint P2P (most of the work) runs on the overlapped step's P2P
    # stream while this stream recomputes the expansions and runs the tree
    # adjoints; the L2P part of the particle gradient goes to its own rows
    # and is added to the P2P rows after the join
    main = torch.cuda.current_stream(dev)
    side = fmm._side_stream(dev)
    side.wait_stream(main)
    with torch.cuda.stream(side):
        _p2p_adjoint(state, code, nat.current_stream(), cot4, grad4)
    lown = _recompute(fmm, state, s, code)
    grad_l4 = torch.zeros((n, 4), dtype=dt, device=dev)
    lbar = torch.zeros_like(state.locals)
    nat.call("jfm_leaf_adjoint_a", s, code, state.order, state.pos4.data_ptr(), cot4.data_ptr(), n, topo.leaf_size,
             topo.num_internal, state.locals.data_ptr(), state.centers4.data_ptr(), state.G, grad_l4.data_ptr(),
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
        if ag is None:
            nat.call("jfm_m2l", s, code, state.order, state.m2l_targets.data_ptr(), nt, state.far_off.data_ptr(),
                     state.far_src.data_ptr(), lbar.data_ptr(), state.centers4.data_ptr(), mbar.data_ptr())
            nat.call("jfm_center_adjoint", s, code, state.order, state.m2l_targets.data_ptr(), nt,
                     state.far_off.data_ptr(), state.far_src.data_ptr(), state.multipoles.data_ptr(),
                     lown.data_ptr(), lbar.data_ptr(), mbar.data_ptr(), state.centers4.data_ptr(), cbar4.data_ptr())
        else:
            for i, g in enumerate(ag.gears):
                if ag.num_targets[i]:
                    nat.call("jfm_m2l_gear", s, code, state.order, g, 1, ag.targets[i].data_ptr(),
                             ag.num_targets[i], ag.offsets[i].data_ptr(), ag.sources[i].data_ptr(), lbar.data_ptr(),
                             state.centers4.data_ptr(), mbar.data_ptr())
            nat.call("jfm_center_adjoint_gears", s, code, state.order, state.m2l_targets.data_ptr(), nt,
                     state.far_off.data_ptr(), state.far_src.data_ptr(), ag.fwd.data_ptr(), ag.rev.data_ptr(),
                     state.multipoles.data_ptr(), lown.data_ptr(), lbar.data_ptr(), mbar.data_ptr(),
                     state.centers4.data_ptr(), cbar4.data_ptr())
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
    main.wait_stream(side)
    grad4.add_(grad_l4)
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
        n = state.num_particles
        acc = torch.empty((n, 3), dtype=state.dtype, device=state.device)
        pot = torch.empty((n,), dtype=state.dtype, device=state.device)
        # refresh + evaluation with the P2P beside the tree passes (the same
        # bits as refresh_prepared_state followed by the one-stream evaluation)
        fmm.step_device(state, x, m, acc, pot)
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
