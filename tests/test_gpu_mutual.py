"""GPU: the momentum-conserving (mutual) lane with block-step level weights.

Mirrors the reference's tests/integration/test_mutual_fmm.py (near-only
round-off case :179-193, force accuracy :201-217, momentum :220-237 and
:281-292, level partition :308-335, cross-rung kick :338-361, uniform-rung
split :364-386, cell-level split :389-425, fused boundary :428-454, rollout
:595-619) on libjfm's kernels, and checks parity with the float64
restatement (oracle/mutual.py) on the same inputs: 1e-12 in float64, 1e-5
in float32.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mutual as om

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SOFT = 1e-2
K_MAX = 3


def _system(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 1.0, (n, 3)), rng.uniform(0.5, 1.5, n)


def _build(n, *, theta, order, leaf, seed=0, dtype=np.float64):
    from jaccpot_amd import mutual as mu

    x, m = _system(n, seed)
    x, m = x.astype(dtype), m.astype(dtype)
    topo, _ = mu.build_mutual_topology(x, m, theta=theta, order=order, leaf_size=leaf)
    st = mu.build_mutual_state(topo, softening=SOFT)
    return x, m, topo, st


def _rungs(n, seed=11):
    return np.random.default_rng(seed).integers(0, K_MAX + 1, n).astype(np.int32)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_near_only_matches_direct_sum_to_round_off(require_gpu):
    from jaccpot_amd import mutual as mu

    print("[gpu] mutual near-only n=512 theta 0.3 leaf 32", flush=True)
    x, m, topo, st = _build(512, theta=0.3, order=4, leaf=32)
    assert topo.num_far_pairs == 0
    acc = mu.mutual_accelerations(st, x, m)
    ref = om.level_direct_sum(x, m, eps=SOFT)
    print(f"[gpu]   rel {_rel(acc, ref):.2e} momentum {om.momentum_residual(acc, m):.2e}", flush=True)
    assert _rel(acc, ref) < 1e-13
    assert om.momentum_residual(acc, m) < 1e-14


@pytest.mark.parametrize("theta, order, tol", [(0.7, 4, 3e-3), (0.7, 6, 3e-3), (1.0, 4, 3e-2), (0.9, 6, 1e-2)])
def test_total_force_within_fmm_tolerance(theta, order, tol, require_gpu):
    from jaccpot_amd import mutual as mu

    print(f"[gpu] mutual force accuracy theta {theta} p {order}", flush=True)
    x, m, topo, st = _build(2048, theta=theta, order=order, leaf=16)
    assert topo.num_far_pairs > 0
    acc = mu.mutual_accelerations(st, x, m)
    err = _rel(acc, om.level_direct_sum(x, m, eps=SOFT))
    print(f"[gpu]   rel err {err:.2e}", flush=True)
    assert 1e-12 < err < tol


@pytest.mark.parametrize("theta", [0.5, 0.7, 1.0])
@pytest.mark.parametrize("order", [2, 4, 6])
def test_momentum_is_exact_independently_of_theta_and_order(theta, order, require_gpu):
    from jaccpot_amd import mutual as mu

    print(f"[gpu] mutual momentum theta {theta} p {order}", flush=True)
    x, m, topo, st = _build(1024, theta=theta, order=order, leaf=16)
    acc = mu.mutual_accelerations(st, x, m)
    res = om.momentum_residual(acc, m)
    print(f"[gpu]   residual {res:.2e}", flush=True)
    assert res < 1e-13


def test_momentum_residual_beats_force_error(require_gpu):
    from jaccpot_amd import mutual as mu

    x, m, topo, st = _build(2048, theta=1.0, order=2, leaf=16)
    assert topo.num_far_pairs > 0
    acc = mu.mutual_accelerations(st, x, m)
    err = _rel(acc, om.level_direct_sum(x, m, eps=SOFT))
    mom = om.momentum_residual(acc, m)
    print(f"[gpu] mutual theta 1.0 p 2: force err {err:.2e} momentum {mom:.2e}", flush=True)
    assert err > 1e-4 and mom < err * 1e-8


@pytest.mark.parametrize("dtype, tol, leaf", [(np.float64, 1e-12, 16), (np.float32, 1e-5, 16), (np.float64, 1e-12, 12),
                                             (np.float64, 1e-12, 24)], ids=["fp64", "fp32", "fp64_leaf12",
                                                                             "fp64_leaf24"])
def test_weighted_matches_cpu_restatement(dtype, tol, leaf, require_gpu):
    """Rung-weighted accelerations and the near/far split against oracle/mutual.py
    (leaves of 12 and 24 take the general slot decode)."""
    from jaccpot_amd import mutual as mu

    n = 3000
    print(f"[gpu] mutual weighted vs restatement n={n} {np.dtype(dtype).name} leaf {leaf}", flush=True)
    x, m, topo, st = _build(n, theta=0.8, order=5, leaf=leaf, seed=3, dtype=dtype)
    rung = _rungs(n, 5)
    w = np.array([0.7, 0.3, 0.2, 0.05])
    got = mu.mutual_weighted_accelerations(st, x, m, rung=rung, level_weights=w, return_parts=True)
    x64, m64 = x.astype(np.float64), m.astype(np.float64)
    acc, near, far = om.mutual_accelerations(x64, m64, leaf_size=leaf, p=5, theta=0.8, eps=SOFT, rung=rung,
                                             level_weights=w, return_parts=True)
    e = (_rel(got.accelerations, acc), _rel(got.near_forces, near), _rel(got.far_forces, far))
    print(f"[gpu]   rel acc {e[0]:.2e} near {e[1]:.2e} far {e[2]:.2e}", flush=True)
    assert max(e) < tol
    full = mu.mutual_accelerations(st, x, m)
    ref = om.mutual_accelerations(x64, m64, leaf_size=leaf, p=5, theta=0.8, eps=SOFT)
    assert _rel(full, ref) < tol


@pytest.mark.parametrize("dtype, leaf", [(np.float64, 16), (np.float32, 16), (np.float64, 12)],
                         ids=["f64", "f32", "f64_leaf12"])
def test_near_work_items_match_per_leaf_launch_and_restatement(dtype, leaf, require_gpu):
    """Near lists cut into work items of two source leaves, so every leaf with
    a longer list is split and its partials go through the fixed-order
    reduction: the near forces equal the one-workgroup-per-leaf launch to the
    working precision's round-off, the accelerations match the restatement,
    momentum stays at round-off, and repeated calls are bit-identical."""
    from jaccpot_amd import mutual as mu
    from jaccpot_amd.nearsplit import build_p2p_items

    n = 3000
    print(f"[gpu] mutual near work items n={n} {np.dtype(dtype).name} leaf {leaf}", flush=True)
    x, m, topo, st = _build(n, theta=0.8, order=5, leaf=leaf, seed=3, dtype=dtype)
    prep = st.prepared
    assert st.near_items is None  # every near list fits one default item at this size
    items = build_p2p_items(np.asarray(prep.topology.near_off), leaf, prep.dtype, prep.device, chunk_slots=2 * leaf)
    assert items is not None and items.num_split > 0 and items.num_items > topo.num_leaves
    rung = _rungs(n, 5)
    w = np.array([0.7, 0.3, 0.2, 0.05])
    per_leaf = mu.mutual_weighted_accelerations(st, x, m, rung=rung, level_weights=w, return_parts=True)
    st.near_items = items
    split = mu.mutual_weighted_accelerations(st, x, m, rung=rung, level_weights=w, return_parts=True)
    again = mu.mutual_weighted_accelerations(st, x, m, rung=rung, level_weights=w, return_parts=True)
    unit = mu.mutual_accelerations(st, x, m)
    st.near_items = None
    x64, m64 = x.astype(np.float64), m.astype(np.float64)
    acc, near, _ = om.mutual_accelerations(x64, m64, leaf_size=leaf, p=5, theta=0.8, eps=SOFT, rung=rung,
                                           level_weights=w, return_parts=True)
    ref_unit = om.mutual_accelerations(x64, m64, leaf_size=leaf, p=5, theta=0.8, eps=SOFT)
    f64 = dtype == np.float64
    d_leaf = _rel(split.near_forces, np.asarray(per_leaf.near_forces, np.float64))
    e = (_rel(split.accelerations, acc), _rel(split.near_forces, near), _rel(unit, ref_unit))
    mom = om.momentum_residual(split.accelerations, m64)
    print(f"[gpu]   items {items.num_items} split leaves {items.num_split}: vs per-leaf {d_leaf:.2e}, "
          f"rel acc {e[0]:.2e} near {e[1]:.2e} unit {e[2]:.2e}, momentum {mom:.2e}", flush=True)
    assert d_leaf < (1e-14 if f64 else 1e-6)
    assert max(e) < (1e-12 if f64 else 1e-5)
    assert mom < (1e-14 if f64 else 1e-6)
    assert np.array_equal(split.near_forces, again.near_forces)


def test_default_state_splits_long_near_lists(require_gpu):
    """A truncated Hernquist sphere (N=20000, leaf 16): its sparse outer leaves
    see every particle, so build_mutual_state itself cuts those near lists
    into work items; the result equals the one-workgroup-per-leaf launch to
    round-off, with momentum at round-off, unit and level weights."""
    from jaccpot_amd import ics
    from jaccpot_amd import mutual as mu

    n = 20000
    print(f"[gpu] mutual default work items, Hernquist n={n} leaf 16", flush=True)
    x, m = ics.galaxy(n, 3, halo_fraction=1.0)
    topo, _ = mu.build_mutual_topology(x, m, theta=0.7, order=4, leaf_size=16)
    st = mu.build_mutual_state(topo, softening=1e-3)
    items = st.near_items
    assert items is not None and items.num_split > 0
    rung = _rungs(n, 7)
    w = np.array([1.0, 0.5, 0.25, 0.125])
    split = mu.mutual_accelerations(st, x, m)
    split_w = mu.mutual_weighted_accelerations(st, x, m, rung=rung, level_weights=w)
    st.near_items = None
    per_leaf = mu.mutual_accelerations(st, x, m)
    per_leaf_w = mu.mutual_weighted_accelerations(st, x, m, rung=rung, level_weights=w)
    st.near_items = items
    d, dw = _rel(split, per_leaf), _rel(split_w, per_leaf_w)
    mom = om.momentum_residual(split, m)
    print(f"[gpu]   items {items.num_items} split leaves {items.num_split}: vs per-leaf {d:.2e} / {dw:.2e}, "
          f"momentum {mom:.2e}", flush=True)
    assert d < 1e-14 and dw < 1e-14 and mom < 1e-14


def test_levels_partition_and_conserve_momentum(require_gpu):
    from jaccpot_amd import mutual as mu

    n = 1024
    print("[gpu] mutual level partition", flush=True)
    x, m, _, st = _build(n, theta=0.9, order=4, leaf=16)
    rung = _rungs(n)
    total = mu.mutual_accelerations(st, x, m)
    levels = [mu.mutual_level_accelerations(st, x, m, rung=rung, level=k, k_max=K_MAX) for k in range(K_MAX + 1)]
    for a in levels:
        assert np.linalg.norm(a) > 0.0
        assert om.momentum_residual(a, m) < 1e-13
    assert _rel(sum(levels), total) < 1e-13


def test_uniform_and_mixed_rung_splits(require_gpu):
    """One rung for all: the split is the dense oracle's (other levels exactly
    zero).  Mixed rungs: the cell-level far split over-refines per level but
    the levels still sum to the total."""
    from jaccpot_amd import mutual as mu

    n = 1024
    x, m, _, st = _build(n, theta=0.9, order=6, leaf=16)
    uni = np.full(n, 2, np.int32)
    for k in range(K_MAX + 1):
        got = mu.mutual_level_accelerations(st, x, m, rung=uni, level=k, k_max=K_MAX)
        if k != 2:
            assert float(np.abs(got).max()) == 0.0
        else:
            assert _rel(got, om.level_direct_sum(x, m, eps=SOFT, rung=uni, level=2)) < 1e-2
    rung = _rungs(n)
    per = [mu.mutual_level_accelerations(st, x, m, rung=rung, level=k, k_max=K_MAX) for k in range(K_MAX + 1)]
    dense = [om.level_direct_sum(x, m, eps=SOFT, rung=rung, level=k) for k in range(K_MAX + 1)]
    assert _rel(sum(per), sum(dense)) < 3e-2
    assert max(_rel(per[k], dense[k]) for k in range(K_MAX + 1)) > 1e-2


def test_fused_boundary_equals_weighted_sum_of_levels(require_gpu):
    from jaccpot_amd import mutual as mu

    n = 1024
    x, m, _, st = _build(n, theta=0.9, order=4, leaf=16)
    rung = _rungs(n)
    per = [mu.mutual_level_accelerations(st, x, m, rung=rung, level=k, k_max=K_MAX) for k in range(K_MAX + 1)]
    for s in range(mu.n_sub(K_MAX) + 1):
        w = mu.boundary_level_weights(s, K_MAX, 0.05)
        fused = mu.mutual_weighted_accelerations(st, x, m, rung=rung, level_weights=w)
        expected = sum(float(w[k]) * per[k] for k in range(K_MAX + 1))
        err = float(np.linalg.norm(fused - expected) / (np.linalg.norm(expected) + 1e-300))
        assert err < 1e-12, f"boundary {s}: {err}"
        assert om.momentum_residual(fused, m) < 1e-13


def test_cross_rung_pair_gets_equal_and_opposite_kick(require_gpu):
    from jaccpot_amd import BlockStepFMM

    x = np.array([[-0.5, 0.0, 0.0], [0.5, 0.2, -0.1]])
    m = np.array([1.0, 3.0])
    rung = np.array([0, 1], np.int32)
    fmm = BlockStepFMM(softening=SOFT, k_max=1, theta=0.6, max_order=4, leaf_size=1)
    fmm.prepare(x, m)
    l0 = fmm.level_accelerations(x, m, rung=rung, level=0)
    l1 = fmm.level_accelerations(x, m, rung=rung, level=1)
    assert float(np.linalg.norm(l0)) == 0.0
    f = m[:, None] * l1
    assert float(np.linalg.norm(f[0] + f[1])) < 1e-15 * float(np.linalg.norm(f[0]))
    with pytest.raises(ValueError):
        fmm.level_accelerations(x, m, rung=np.array([0, 2]), level=0)


def test_zero_mass_particles_get_zero_acceleration(require_gpu):
    from jaccpot_amd import mutual as mu

    x, m = _system(600, 8)
    m[::7] = 0.0
    topo, _ = mu.build_mutual_topology(x, m, theta=0.8, order=4, leaf_size=16)
    st = mu.build_mutual_state(topo, softening=SOFT)
    acc = mu.mutual_accelerations(st, x, m)
    assert np.all(acc[::7] == 0.0) and np.all(np.isfinite(acc))
    ref = om.mutual_accelerations(x, m, leaf_size=16, p=4, theta=0.8, eps=SOFT)
    assert _rel(acc, ref) < 1e-12


def test_block_step_rollout_conserves_momentum_and_bounds_energy(require_gpu):
    from jaccpot_amd import BlockStepFMM
    from oracle.direct import direct_sum

    n = 512
    print("[gpu] mutual block-step rollout n=512 k_max 2", flush=True)
    x, m = _system(n, 21)
    x, m = 0.5 * x, m / n
    v = np.random.default_rng(22).normal(0.0, 0.3, (n, 3))
    fmm = BlockStepFMM(softening=0.05, k_max=2, theta=0.7, max_order=4, leaf_size=16)

    def energy(x, v):
        _, pot = direct_sum(x, m, eps=0.05, return_potential=True)
        return 0.5 * float((m * (v * v).sum(1)).sum()) + 0.5 * float((m * pot).sum())

    p0 = (m[:, None] * v).sum(0)
    scale = float(np.abs(m[:, None] * v).sum())
    e0 = energy(x, v)
    for step in range(4):
        fmm.prepare(x, m)
        acc = fmm.total_accelerations(x, m)
        amag = np.linalg.norm(acc, axis=1)
        rung = np.clip(np.floor(np.log2(np.maximum(amag / np.median(amag), 1.0))), 0, 2).astype(np.int32)
        x, v, _ = fmm.advance_base_step(x, v, m, rung=rung, dt_max=0.01)
    drift = float(np.linalg.norm((m[:, None] * v).sum(0) - p0)) / scale
    de = abs(energy(x, v) - e0) / abs(e0)
    print(f"[gpu]   momentum drift {drift:.2e} energy change {de:.2e}", flush=True)
    assert drift < 1e-13 and de < 5e-3


def test_fp32_momentum_at_single_precision_round_off(require_gpu):
    from jaccpot_amd import mutual as mu

    x, m, topo, st = _build(4096, theta=0.8, order=4, leaf=16, dtype=np.float32)
    rung = _rungs(4096)
    acc = mu.mutual_weighted_accelerations(st, x, m, rung=rung, level_weights=np.array([1.0, 0.5, 0.25, 0.125]))
    res = om.momentum_residual(acc, m.astype(np.float64))
    print(f"[gpu] mutual fp32 momentum residual {res:.2e}", flush=True)
    assert acc.dtype == np.float32 and res < 1e-6


@pytest.mark.parametrize("n, leaf", [(1, 16), (5, 16), (40, 64)], ids=["one", "single_leaf", "below_leaf"])
def test_small_systems(n, leaf, require_gpu):
    """A single particle feels nothing; one leaf is the direct sum."""
    from jaccpot_amd import mutual as mu

    x, m = _system(n, 30 + n)
    topo, _ = mu.build_mutual_topology(x, m, theta=0.6, order=3, leaf_size=leaf)
    assert topo.num_far_pairs == 0 and topo.num_near_pairs == 0
    st = mu.build_mutual_state(topo, softening=SOFT)
    acc = mu.mutual_accelerations(st, x, m)
    if n == 1:
        assert np.all(acc == 0.0)
    else:
        assert _rel(acc, om.level_direct_sum(x, m, eps=SOFT)) < 1e-13
        assert om.momentum_residual(acc, m) < 1e-15


def _mutual_loss_grad(mu, st, x, m, w, rung, lw):
    xt = torch.as_tensor(x, device="cuda").requires_grad_(True)
    mt = torch.as_tensor(m, device="cuda").requires_grad_(True)
    acc = mu.mutual_weighted_accelerations(st, xt, mt, rung=rung, level_weights=lw)
    (acc * torch.as_tensor(w, device="cuda")).sum().backward()
    return xt.grad.double().cpu().numpy(), mt.grad.double().cpu().numpy()


def _mutual_loss(mu, st, x, m, w, rung, lw):
    a = mu.mutual_weighted_accelerations(st, x, m, rung=rung, level_weights=lw)
    return float((np.asarray(a, np.float64) * w).sum())


@pytest.mark.parametrize("weighted", [False, True], ids=["unit", "levels"])
def test_gradient_matches_central_differences(weighted, require_gpu):
    """Reverse mode of the mutual accelerations (fixed topology, rungs and
    weights) against float64 central differences."""
    from jaccpot_amd import mutual as mu

    n = 1500
    print(f"[gpu] mutual gradient fd ({'levels' if weighted else 'unit weights'})", flush=True)
    x, m, topo, st = _build(n, theta=0.8, order=5, leaf=16, seed=12)
    assert topo.num_far_pairs > 0
    rng = np.random.default_rng(13)
    w = rng.normal(size=(n, 3))
    rung = _rungs(n, 14) if weighted else None
    lw = np.array([1.0, 0.5, 0.25, 0.125]) if weighted else None
    gx, gm = _mutual_loss_grad(mu, st, x, m, w, rung, lw)
    fds, ans = [], []
    for k in range(6):
        vx = rng.normal(size=x.shape)
        vm = rng.normal(size=m.shape) * np.abs(m).mean()
        h = 1e-6
        fd = (_mutual_loss(mu, st, x + h * vx, m + h * vm, w, rung, lw)
              - _mutual_loss(mu, st, x - h * vx, m - h * vm, w, rung, lw)) / (2 * h)
        an = float((gx * vx).sum() + (gm * vm).sum())
        print(f"[gpu]   dir {k}: fd {fd:.10e} adjoint {an:.10e} rel {abs(fd - an) / max(abs(an), 1e-30):.2e}",
              flush=True)
        fds.append(fd)
        ans.append(an)
    rms = float(np.sqrt(np.mean(np.square(ans))))
    for fd, an in zip(fds, ans):
        assert abs(fd - an) <= 1e-6 * max(abs(an), rms)


@pytest.mark.parametrize("weighted", [False, True], ids=["unit", "levels"])
def test_gradient_through_near_work_items_equals_per_leaf_gradient(weighted, require_gpu):
    """Near lists cut into work items of two source leaves in both passes (the
    adjoint's split leaves go through the adjoint reduction): the gradient
    equals the one-workgroup-per-leaf gradient to float64 round-off."""
    from jaccpot_amd import mutual as mu
    from jaccpot_amd.nearsplit import build_p2p_items

    n = 1500
    print(f"[gpu] mutual gradient through near work items ({'levels' if weighted else 'unit weights'})", flush=True)
    x, m, topo, st = _build(n, theta=0.8, order=5, leaf=16, seed=12)
    prep = st.prepared
    items = build_p2p_items(np.asarray(prep.topology.near_off), 16, prep.dtype, prep.device, chunk_slots=32)
    assert items is not None and items.num_split > 0
    w = np.random.default_rng(13).normal(size=(n, 3))
    rung = _rungs(n, 14) if weighted else None
    lw = np.array([1.0, 0.5, 0.25, 0.125]) if weighted else None
    gx0, gm0 = _mutual_loss_grad(mu, st, x, m, w, rung, lw)
    st.near_items = items
    gx1, gm1 = _mutual_loss_grad(mu, st, x, m, w, rung, lw)
    st.near_items = None
    ex, em = _rel(gx1, gx0), _rel(gm1, gm0)
    print(f"[gpu]   items {items.num_items}, split leaves {items.num_split}: positions {ex:.2e} masses {em:.2e}",
          flush=True)
    assert ex < 1e-13 and em < 1e-13


def test_unit_weight_gradient_equals_fixed_order_gradient(require_gpu):
    """With unit weights the mutual acceleration is the plain FMM's on the same
    topology, so the two reverse passes agree to rounding."""
    from jaccpot_amd import mutual as mu

    n = 2000
    x, m, topo, st = _build(n, theta=0.7, order=4, leaf=16, seed=15)
    fmm_state = topo.prepared
    w = np.random.default_rng(16).normal(size=(n, 3))
    gx, gm = _mutual_loss_grad(mu, st, x, m, w, None, None)
    from jaccpot_amd import FastMultipoleMethod

    fmm = FastMultipoleMethod(precision="fp64", theta=0.7, softening=SOFT)
    fmm_state.softening = SOFT
    xt = torch.as_tensor(x, device="cuda").requires_grad_(True)
    mt = torch.as_tensor(m, device="cuda").requires_grad_(True)
    acc = fmm.differentiable_accelerations(fmm_state, xt, mt)
    (acc * torch.as_tensor(w, device="cuda")).sum().backward()
    ex = float(np.linalg.norm(gx - xt.grad.cpu().numpy()) / np.linalg.norm(gx))
    em = float(np.linalg.norm(gm - mt.grad.cpu().numpy()) / np.linalg.norm(gm))
    print(f"[gpu] mutual vs fixed-order gradient: positions {ex:.2e} masses {em:.2e}", flush=True)
    assert ex < 1e-11 and em < 1e-11
