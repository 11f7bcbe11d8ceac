"""GPU: exact FMM gradients (adjoint kernels) against central finite differences
at the frozen topology, and against the direct-sum gradient."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import ics
from oracle.direct import direct_sum_vjp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _loss_and_grad(fmm, state, x, m, w, u):
    from jaccpot_amd.autodiff import fmm_vjp

    xd = torch.as_tensor(x, device="cuda").to(state.dtype)
    md = torch.as_tensor(m, device="cuda").to(state.dtype)
    wd = torch.as_tensor(w, device="cuda").to(state.dtype)
    ud = torch.as_tensor(u, device="cuda").to(state.dtype)
    gx, gm = fmm_vjp(fmm, state, xd, md, wd, ud)
    return gx.cpu().numpy().astype(np.float64), gm.cpu().numpy().astype(np.float64)


def _loss(fmm, state, x, m, w, u):
    xd = torch.as_tensor(x, device="cuda").to(state.dtype)
    md = torch.as_tensor(m, device="cuda").to(state.dtype)
    fmm.refresh_prepared_state(state, xd, md)
    n = x.shape[0]
    acc = torch.empty((n, 3), dtype=state.dtype, device="cuda")
    pot = torch.empty((n,), dtype=state.dtype, device="cuda")
    fmm._evaluate_device(state, acc, pot)
    return float((acc.double().cpu().numpy() * w).sum() + (pot.double().cpu().numpy() * u).sum())


CASES = [
    # name, gen, n, leaf, p, theta, eps
    ("near_only_leaf64", "uniform", 300, 64, 4, 0.6, 1e-2),
    ("plummer1200_p4", "plummer", 1200, 16, 4, 0.6, 1e-2),
    ("uniform2000_p3_eps0", "uniform", 2000, 8, 3, 0.5, 0.0),
    ("blobs1500_p6", "blobs", 1500, 16, 6, 0.7, 1e-2),
    ("plummer1500_p7", "plummer", 1500, 16, 7, 0.75, 1e-2),
    ("plummer1500_p8", "plummer", 1500, 16, 8, 0.75, 1e-2),
    # equal masses on a jittered-free lattice: many node centres share x and y
    # (far-pair separations along the z axis, where the M2L alignment angle is
    # fixed by convention and has no derivative)
    ("lattice1728_p4", "lattice", 1728, 8, 4, 0.6, 1e-2),
]


def _lattice(n, seed):
    k = round(n ** (1 / 3))
    assert k ** 3 == n
    g = (np.arange(k) + 0.5) / k
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return x, np.full(n, 1.0 / n)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp64_gradient_matches_central_differences(case, require_gpu):
    from jaccpot_amd import FastMultipoleMethod

    name, gen, n, leaf, p, theta, eps = case
    print(f"[gpu] adjoint fd {name}", flush=True)
    x, m = _lattice(n, 41) if gen == "lattice" else getattr(ics, gen)(n, 41)
    rng = np.random.default_rng(7)
    w = rng.normal(size=(n, 3))
    u = rng.normal(size=n)
    fmm = FastMultipoleMethod(precision="fp64", theta=theta, softening=eps)
    state = fmm.prepare_state(x, m, leaf_size=leaf, max_order=p)
    gx, gm = _loss_and_grad(fmm, state, x, m, w, u)
    fds, ans = [], []
    for k in range(6):
        vx = rng.normal(size=x.shape)
        vm = rng.normal(size=m.shape) * np.abs(m).mean()
        h = 1e-6
        fd = (_loss(fmm, state, x + h * vx, m + h * vm, w, u) - _loss(fmm, state, x - h * vx, m - h * vm, w, u)) / (
            2 * h)
        an = float((gx * vx).sum() + (gm * vm).sum())
        print(f"[gpu]   dir {k}: fd {fd:.10e} adjoint {an:.10e} rel {abs(fd - an) / max(abs(an), 1e-30):.2e}",
              flush=True)
        fds.append(fd)
        ans.append(an)
    # per direction, relative to the larger of the directional derivative and
    # the RMS over directions (central differences lose digits on directions
    # whose derivative is small next to the curvature of unsoftened pairs)
    rms = float(np.sqrt(np.mean(np.square(ans))))
    for fd, an in zip(fds, ans):
        assert abs(fd - an) <= 1e-6 * max(abs(an), rms)


def test_gradient_close_to_direct_sum_gradient(require_gpu):
    """At N=4096 the FMM gradient approaches the direct-sum gradient (truncation level)."""
    from jaccpot_amd import FastMultipoleMethod

    x, m = ics.blobs(4096, 4)
    rng = np.random.default_rng(8)
    tgt = rng.choice(4096, 1024, replace=False)
    eps = 1e-2
    fmm = FastMultipoleMethod(precision="fp64", theta=0.5, softening=eps)
    state = fmm.prepare_state(x, m, leaf_size=32, max_order=6)
    acc, pot = fmm.evaluate_prepared_state(state, return_potential=True)
    # C5-style loss over 1024 targets: sum_t m_t phi_t + 0.5 sum_t |a_t|^2
    w = np.zeros((4096, 3))
    u = np.zeros(4096)
    w[tgt] = acc[tgt]
    u[tgt] = m[tgt]
    gx, gm = _loss_and_grad(fmm, state, x, m, w, u)
    # the phi term's explicit mass dependence: d/dm_t (m_t phi_t) = phi_t
    gm[tgt] += pot[tgt]
    from oracle.direct import direct_sum

    ad, pd = direct_sum(x, m, eps=eps, return_potential=True)
    wd = np.zeros((4096, 3))
    ud = np.zeros(4096)
    wd[tgt] = ad[tgt]
    ud[tgt] = m[tgt]
    rx, rm = direct_sum_vjp(x, m, wd, ud, eps=eps)
    rm[tgt] += pd[tgt]
    ex = np.linalg.norm(gx - rx) / np.linalg.norm(rx)
    em = np.linalg.norm(gm - rm) / np.linalg.norm(rm)
    print(f"[gpu] grad vs direct-sum grad: positions {ex:.3e} masses {em:.3e}", flush=True)
    assert ex < 1e-3 and em < 1e-3


def test_autograd_function_and_fp32(require_gpu):
    from jaccpot_amd import FastMultipoleMethod

    x, m = ics.plummer(3000, 9)
    fmm32 = FastMultipoleMethod(precision="fp32", theta=0.6, softening=1e-2)
    st32 = fmm32.prepare_state(x.astype(np.float32), m.astype(np.float32), leaf_size=16, max_order=4)
    xt = torch.as_tensor(x, dtype=torch.float32, device="cuda").requires_grad_(True)
    mt = torch.as_tensor(m, dtype=torch.float32, device="cuda").requires_grad_(True)
    acc, pot = fmm32.differentiable_accelerations_and_potential(st32, xt, mt)
    loss = (mt * pot).sum() + 0.5 * (acc * acc).sum()
    loss.backward()
    fmm64 = FastMultipoleMethod(precision="fp64", theta=0.6, softening=1e-2)
    st64 = fmm64.prepare_state(x, m, leaf_size=16, max_order=4)
    x64 = torch.as_tensor(x, device="cuda").requires_grad_(True)
    m64 = torch.as_tensor(m, device="cuda").requires_grad_(True)
    a64, p64 = fmm64.differentiable_accelerations_and_potential(st64, x64, m64)
    l64 = (m64 * p64).sum() + 0.5 * (a64 * a64).sum()
    l64.backward()
    ex = (torch.linalg.norm(xt.grad.double() - x64.grad) / torch.linalg.norm(x64.grad)).item()
    em = (torch.linalg.norm(mt.grad.double() - m64.grad) / torch.linalg.norm(m64.grad)).item()
    print(f"[gpu] fp32 vs fp64 gradient: positions {ex:.3e} masses {em:.3e}", flush=True)
    assert ex < 1e-4 and em < 1e-4


@pytest.mark.parametrize("eps", [1e-2, 0.0], ids=["softened", "eps0"])
def test_split_near_lists_give_the_same_gradient(eps, require_gpu):
    """The adjoint P2P over work items (long near lists cut into slot ranges,
    partials added in item order) gives the one-workgroup-per-leaf gradient
    to float64 rounding."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd.nearsplit import build_p2p_items

    print(f"[gpu] adjoint split near lists eps={eps}", flush=True)
    n = 3000
    x, m = ics.plummer(n, 43)
    rng = np.random.default_rng(9)
    w = rng.normal(size=(n, 3))
    u = rng.normal(size=n)
    fmm = FastMultipoleMethod(precision="fp64", theta=0.6, softening=eps, split_near=False)
    state = fmm.prepare_state(x, m, leaf_size=16, max_order=4)
    gx0, gm0 = _loss_and_grad(fmm, state, x, m, w, u)
    items = build_p2p_items(state.topology.near_off, 16, state.dtype, "cuda", chunk_slots=64)
    assert items.num_split > 50
    state.extra["p2p_items"] = items
    gx1, gm1 = _loss_and_grad(fmm, state, x, m, w, u)
    assert np.linalg.norm(gx1 - gx0) / np.linalg.norm(gx0) < 1e-13
    assert np.linalg.norm(gm1 - gm0) / np.linalg.norm(gm0) < 1e-13


def test_per_pair_gear_gradient_matches_central_differences(require_gpu):
    """Per-pair M2L orders (gears 2, 4, 6 drawn at random per far pair, so a
    pair and its mirror mostly run at different orders): the adjoint of the
    gear M2L and of its centre terms against float64 central differences at
    the frozen topology and gears."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd.autodiff import adjoint_gears
    from jaccpot_amd.learned import build_gear_schedule

    n = 1500
    print("[gpu] adjoint fd, random per-pair gears (2,4,6), plummer 1500", flush=True)
    x, m = ics.plummer(n, 44)
    rng = np.random.default_rng(10)
    w = rng.normal(size=(n, 3))
    u = rng.normal(size=n)
    fmm = FastMultipoleMethod(precision="fp64", theta=0.7, softening=1e-2)
    state = fmm.prepare_state(x, m, leaf_size=16, max_order=6)
    K = state.topology.num_far_pairs
    g = rng.choice(np.array([2, 4, 6], np.int8), size=K)
    state.extra["gears"] = build_gear_schedule(state, g, (2, 4, 6))
    state.extra["gear_of_pair"] = g
    ag = adjoint_gears(state)
    fwd, rev = ag.fwd.cpu().numpy(), ag.rev.cpu().numpy()
    assert np.array_equal(fwd, g) and (fwd != rev).mean() > 0.5
    # the mirror of the mirror is the pair itself
    assert np.bincount(rev, minlength=7)[[2, 4, 6]].tolist() == np.bincount(g, minlength=7)[[2, 4, 6]].tolist()
    gx, gm = _loss_and_grad(fmm, state, x, m, w, u)
    fds, ans = [], []
    for k in range(6):
        vx = rng.normal(size=x.shape)
        vm = rng.normal(size=m.shape) * np.abs(m).mean()
        h = 1e-6
        fd = (_loss(fmm, state, x + h * vx, m + h * vm, w, u) - _loss(fmm, state, x - h * vx, m - h * vm, w, u)) / (
            2 * h)
        an = float((gx * vx).sum() + (gm * vm).sum())
        print(f"[gpu]   dir {k}: fd {fd:.10e} adjoint {an:.10e} rel {abs(fd - an) / max(abs(an), 1e-30):.2e}",
              flush=True)
        fds.append(fd)
        ans.append(an)
    rms = float(np.sqrt(np.mean(np.square(ans))))
    for fd, an in zip(fds, ans):
        assert abs(fd - an) <= 1e-6 * max(abs(an), rms)
    # the gears change the gradient: the all-top-gear gradient is a different one
    state.extra.pop("gears")
    state.extra.pop("gear_of_pair")
    gx6, _ = _loss_and_grad(fmm, state, x, m, w, u)
    assert np.linalg.norm(gx6 - gx) > 1e-8 * np.linalg.norm(gx)


LEARNED_GRAD_CASES = [
    # name, generator, gears, leaf
    ("galaxy16384_gears246", "galaxy", (2, 4, 6), 16),
    ("plummer16384_gears368", "plummer", (3, 6, 8), 8),
]


@pytest.mark.parametrize("case", LEARNED_GRAD_CASES, ids=[c[0] for c in LEARNED_GRAD_CASES])
def test_learned_mixed_fp32_gradient_matches_fp64(case, require_gpu):
    """A shipped learned model with float64 top levels in float32: autograd
    gradients of the C5-style loss against the float64 gradient at the same
    topology and the same per-pair gears."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd.learned import build_gear_schedule

    name, gen, gears, leaf = case
    n = 16384
    print(f"[gpu] learned+mixed fp32 gradient vs fp64, {name}", flush=True)
    x, m = getattr(ics, gen)(n, 5)
    x32, m32 = x.astype(np.float32), m.astype(np.float32)
    fmm32 = FastMultipoleMethod(precision="fp32", theta=0.6, softening=1e-3, adaptive_order=True, p_gears=gears,
                                gear_tolerance=-1.5, mixed_precision=True, order_judge_targets=0)
    st32 = fmm32.prepare_state(x32, m32, leaf_size=leaf, max_order=gears[-1])
    assert st32.extra.get("mixed") is not None
    g = st32.extra["gear_of_pair"]
    counts = {int(k): int(np.sum(g == k)) for k in gears}
    print(f"[gpu]   gears {counts}", flush=True)
    assert min(counts.values()) >= 0.005 * g.size
    xt = torch.as_tensor(x32, device="cuda").requires_grad_(True)
    mt = torch.as_tensor(m32, device="cuda").requires_grad_(True)
    acc, pot = fmm32.differentiable_accelerations_and_potential(st32, xt, mt)
    loss = (mt * pot)[:1024].sum() + 0.5 * (acc[:1024] * acc[:1024]).sum()
    loss.backward()

    fmm64 = FastMultipoleMethod(precision="fp64", theta=0.6, softening=1e-3)
    st64 = fmm64.prepare_state(x32.astype(np.float64), m32.astype(np.float64), leaf_size=leaf, max_order=gears[-1])
    t32, t64 = st32.topology, st64.topology
    assert np.array_equal(t32.far_off, t64.far_off) and np.array_equal(t32.far_src, t64.far_src)
    st64.extra["gears"] = build_gear_schedule(st64, g, gears)
    st64.extra["gear_of_pair"] = g
    x64 = torch.as_tensor(x32.astype(np.float64), device="cuda").requires_grad_(True)
    m64 = torch.as_tensor(m32.astype(np.float64), device="cuda").requires_grad_(True)
    a64, p64 = fmm64.differentiable_accelerations_and_potential(st64, x64, m64)
    l64 = (m64 * p64)[:1024].sum() + 0.5 * (a64[:1024] * a64[:1024]).sum()
    l64.backward()
    ex = (torch.linalg.norm(xt.grad.double() - x64.grad) / torch.linalg.norm(x64.grad)).item()
    em = (torch.linalg.norm(mt.grad.double() - m64.grad) / torch.linalg.norm(m64.grad)).item()
    print(f"[gpu]   fp32 learned+mixed vs fp64 same gears: positions {ex:.3e} masses {em:.3e}", flush=True)
    assert ex < 1e-4 and em < 1e-4
