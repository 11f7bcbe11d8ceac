"""GPU parity: every HIP kernel against the float64 oracle on the same inputs.

Float64 kernels must match the oracle to 1e-12 (relative); float32 kernels to
1e-5 (relative L2 over all particles, plus a per-particle median bound).  The
topology is built independently on both sides and must agree bit-exactly (the
CPU suite checks that too), so differences here are kernel arithmetic only.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fmm as ofmm
from oracle import harmonics as H
from oracle import ics
from oracle import topology as OT
from oracle.direct import direct_sum, relative_errors

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _fmm(precision, theta, eps, compensated=True):
    from jaccpot_amd import FastMultipoleMethod

    return FastMultipoleMethod(precision=precision, theta=theta, softening=eps, compensated=compensated)


def _rel_rows(a, b):
    """max over rows of |a-b| / max|b| (row scale), robust to zero rows."""
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    return float((np.abs(a - b) / scale).max())


CASES64 = [
    # (name, generator, n, leaf, p, theta, eps)
    ("uniform4096_leaf16_p4_eps0", "uniform", 4096, 16, 4, 0.6, 0.0),
    ("uniform1500_leaf8_p6_eps1e-2", "uniform", 1500, 8, 6, 0.5, 1e-2),
    ("plummer3000_leaf32_p3", "plummer", 3000, 32, 3, 0.7, 1e-3),
    ("blobs2000_leaf64_p8", "blobs", 2000, 64, 8, 0.6, 1e-3),
    ("ragged7_leaf4_p2", "uniform", 7, 4, 2, 0.6, 0.0),
    ("single_particle", "uniform", 1, 16, 4, 0.6, 0.0),
    ("uniform777_leaf5_p5", "uniform", 777, 5, 5, 0.55, 1e-3),
]


def _inputs(gen, n, seed=0):
    return getattr(ics, gen)(n, seed)


@pytest.mark.parametrize("case", CASES64, ids=[c[0] for c in CASES64])
def test_fp64_pipeline_matches_oracle(case, require_gpu):
    name, gen, n, leaf, p, theta, eps = case
    print(f"[gpu] fp64 pipeline {name}", flush=True)
    x, m = _inputs(gen, n)
    fmm = _fmm("fp64", theta, eps)
    state = fmm.prepare_state(x, m, leaf_size=leaf, max_order=p)
    acc, pot = fmm.evaluate_prepared_state(state, return_potential=True)
    torch.cuda.synchronize()

    topo = OT.build_topology(x, m, leaf, theta)
    assert np.array_equal(topo.perm, state.topology.perm)
    assert np.array_equal(topo.far_src, state.topology.far_src)
    xs, ms = x[topo.perm], m[topo.perm]
    C = H.sh_size(p)
    M = ofmm.upward(topo, xs, ms, p)
    Lc = ofmm.downward(topo, M, p)
    acc_s, pot_s = ofmm.evaluate(topo, xs, ms, Lc, p, 1.0, eps)

    centers = state.centers4.cpu().numpy()
    assert np.max(np.abs(centers[:, :3] - topo.center)) <= 1e-13 * max(1.0, np.abs(topo.center).max())
    np.testing.assert_allclose(centers[:, 3], topo.mass, rtol=1e-13)
    gM = state.multipoles.cpu().numpy()[:, :C]
    assert _rel_rows(gM, M) < 1e-12
    gL = state.locals.cpu().numpy()[:, :C]
    if np.abs(Lc).max() > 0:
        assert _rel_rows(gL, Lc) < 1e-12
    ref_acc = np.empty_like(acc_s)
    ref_pot = np.empty_like(pot_s)
    ref_acc[topo.perm] = acc_s
    ref_pot[topo.perm] = pot_s
    if n > 1:
        rel = np.linalg.norm(acc - ref_acc) / np.linalg.norm(ref_acc)
        assert rel < 1e-12, rel
        assert np.median(relative_errors(acc, ref_acc)) < 1e-12
        assert np.linalg.norm(pot - ref_pot) / np.linalg.norm(ref_pot) < 1e-12
    else:
        assert np.all(acc == 0.0) and np.all(pot == 0.0)


CASES32 = [
    ("uniform4096_leaf16_p4", "uniform", 4096, 16, 4, 0.6, 0.0),
    ("plummer8192_leaf32_p4", "plummer", 8192, 32, 4, 0.6, 1e-3),
    ("plummer8192_leaf64_p6", "plummer", 8192, 64, 6, 0.6, 1e-3),
    ("blobs4096_leaf32_p4", "blobs", 4096, 32, 4, 0.6, 1e-3),
]


@pytest.mark.parametrize("case", CASES32, ids=[c[0] for c in CASES32])
def test_fp32_pipeline_matches_oracle(case, require_gpu):
    name, gen, n, leaf, p, theta, eps = case
    print(f"[gpu] fp32 pipeline {name}", flush=True)
    x, m = _inputs(gen, n, seed=1)
    fmm = _fmm("fp32", theta, eps)
    acc = fmm.compute_accelerations(x.astype(np.float32), m.astype(np.float32), leaf_size=leaf, max_order=p)
    torch.cuda.synchronize()
    ref = ofmm.fmm_accelerations(x.astype(np.float32).astype(np.float64), m.astype(np.float32).astype(np.float64),
                                 leaf_size=leaf, p=p, theta=theta, eps=eps)
    rel = np.linalg.norm(acc - ref) / np.linalg.norm(ref)
    errs = relative_errors(acc, ref)
    print(f"[gpu]   rel-L2 {rel:.3e} median {np.median(errs):.3e} max {errs.max():.3e}", flush=True)
    assert rel < 1e-5
    assert np.median(errs) < 1e-5


def test_near_and_far_parts_separately(require_gpu):
    """P2P-only and L2P-only evaluations against the oracle's split."""
    from jaccpot_amd import _native as nat

    x, m = ics.plummer(3000, 5)
    p, leaf, theta, eps = 4, 16, 0.6, 1e-3
    fmm = _fmm("fp64", theta, eps)
    state = fmm.prepare_state(x, m, leaf_size=leaf, max_order=p)
    n = x.shape[0]
    out = {}
    for name, flags in (("near", nat.EVAL_NEAR), ("far", nat.EVAL_FAR)):
        acc = torch.empty((n, 3), dtype=torch.float64, device=state.device)
        pot = torch.empty((n,), dtype=torch.float64, device=state.device)
        fmm._evaluate_device(state, acc, pot, flags=flags)
        out[name] = (acc.cpu().numpy(), pot.cpu().numpy())
    topo = OT.build_topology(x, m, leaf, theta)
    xs, ms = x[topo.perm], m[topo.perm]
    M = ofmm.upward(topo, xs, ms, p)
    Lc = ofmm.downward(topo, M, p)
    zero = np.zeros_like(Lc)
    near_s, near_pot_s = ofmm.evaluate(topo, xs, ms, zero, p, 1.0, eps)
    full_s, full_pot_s = ofmm.evaluate(topo, xs, ms, Lc, p, 1.0, eps)
    far_s, far_pot_s = full_s - near_s, full_pot_s - near_pot_s
    for (g_acc, g_pot), r_acc, r_pot in ((out["near"], near_s, near_pot_s), (out["far"], far_s, far_pot_s)):
        ra = np.empty_like(r_acc)
        rp = np.empty_like(r_pot)
        ra[topo.perm] = r_acc
        rp[topo.perm] = r_pot
        assert np.linalg.norm(g_acc - ra) / np.linalg.norm(ra) < 1e-12
        assert np.linalg.norm(g_pot - rp) / np.linalg.norm(rp) < 1e-12


def test_direct_sum_kernel_matches_oracle(require_gpu):
    from jaccpot_amd import direct_sum_gravitational_acceleration

    x, m = ics.blobs(3000, 7)
    tgt = np.random.default_rng(12345).choice(3000, 257, replace=False)
    for eps in (0.0, 1e-2):
        a, phi = direct_sum_gravitational_acceleration(x, m, softening=eps, target_indices=tgt, return_potential=True)
        ra, rp = direct_sum(x, m, eps=eps, targets=tgt, return_potential=True)
        np.testing.assert_allclose(a.cpu().numpy(), ra, rtol=1e-12, atol=1e-12 * np.abs(ra).max())
        np.testing.assert_allclose(phi.cpu().numpy(), rp, rtol=1e-12)


def test_target_indices_and_input_types(require_gpu):
    x, m = ics.uniform(2048, 9)
    fmm = _fmm("fp64", 0.6, 1e-3)
    full = fmm.compute_accelerations(x, m, leaf_size=16, max_order=4)
    idx = np.array([5, 0, 2047, 5, 100])
    sub, pot = fmm.compute_accelerations(x, m, leaf_size=16, max_order=4, target_indices=idx, return_potential=True)
    np.testing.assert_array_equal(sub, full[idx])
    assert sub.dtype == np.float64 and pot.shape == (5,)
    xt = torch.as_tensor(x, device="cuda")
    mt = torch.as_tensor(m, device="cuda")
    at = fmm.compute_accelerations(xt, mt, leaf_size=16, max_order=4)
    assert isinstance(at, torch.Tensor) and at.device.type == "cuda"
    np.testing.assert_array_equal(at.cpu().numpy(), full)


def test_mixed_precision_top_levels(require_gpu):
    """float32 with float64 top-three-level M2M/L2L matches the oracle to 1e-5 and
    its float64 shadows match the oracle's top-level expansions to float32 input rounding."""
    from jaccpot_amd import FastMultipoleMethod

    x, m = ics.plummer(8192, 11)
    x32, m32 = x.astype(np.float32), m.astype(np.float32)
    xr, mr = x32.astype(np.float64), m32.astype(np.float64)
    p, leaf, theta, eps = 6, 64, 0.6, 1e-3
    fmm = FastMultipoleMethod(precision="fp32", theta=theta, softening=eps, mixed_precision=True)
    state = fmm.prepare_state(x32, m32, leaf_size=leaf, max_order=p)
    acc = fmm.evaluate_prepared_state(state)
    torch.cuda.synchronize()
    topo = OT.build_topology(xr, mr, leaf, theta)
    M = ofmm.upward(topo, xr[topo.perm], mr[topo.perm], p)
    Lc = ofmm.downward(topo, M, p)
    mixed = state.extra["mixed"]
    top = mixed["top"].cpu().numpy()[: mixed["ntop"]]
    C = H.sh_size(p)
    assert mixed["ntop"] == int(np.sum(topo.depth[: topo.num_internal] <= 2))
    m64 = mixed["mult64"].cpu().numpy()[: len(top), :C]
    assert _rel_rows(m64, M[top]) < 2e-6
    l64 = mixed["loc64"].cpu().numpy()[: len(top), :C]
    nz = np.abs(Lc[top]).max(axis=1) > 0
    if nz.any():
        assert _rel_rows(l64[nz], Lc[top][nz]) < 2e-5
    ref = ofmm.fmm_accelerations(xr, mr, leaf_size=leaf, p=p, theta=theta, eps=eps, topo=topo)
    rel = np.linalg.norm(acc - ref) / np.linalg.norm(ref)
    assert rel < 1e-5, rel


@pytest.mark.parametrize("dtype_name", ["fp64", "fp32"])
def test_flat_m2l_matches_per_target_m2l(dtype_name, require_gpu):
    """Lane-per-pair M2L (with targets spanning several 64-pair blocks) equals the
    wave-per-target kernel, and repeats bit for bit."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd import _native as nat
    from jaccpot_amd.flatlist import build_flat_list, launch_flat

    print(f"[gpu] flat m2l {dtype_name}", flush=True)
    x, m = ics.plummer(6000, 17)
    if dtype_name == "fp32":
        x, m = x.astype(np.float32), m.astype(np.float32)
    fmm = FastMultipoleMethod(precision=dtype_name, theta=0.5, softening=1e-3)
    state = fmm.prepare_state(x, m, leaf_size=4, max_order=5)
    topo = state.topology
    counts = np.diff(topo.far_off)
    tgt = np.repeat(np.arange(topo.num_nodes, dtype=np.int32), counts)
    fl = build_flat_list(tgt, topo.far_src, 5, state.dtype, state.device)
    print(f"[gpu]   {topo.num_far_pairs} pairs, max per target {counts.max()}, {fl.num_span} spanning targets",
          flush=True)
    assert fl.num_span > 0 and counts.max() > 128
    s = nat.current_stream()
    code = nat.dtype_code(state.dtype)
    ref = torch.zeros_like(state.locals)
    nat.call("jfm_m2l", s, code, 5, state.m2l_targets.data_ptr(), int(topo.m2l_targets.size), state.far_off.data_ptr(),
             state.far_src.data_ptr(), state.multipoles.data_ptr(), state.centers4.data_ptr(), ref.data_ptr())
    outs = []
    for _ in range(2):
        loc = torch.zeros_like(state.locals)
        launch_flat(fl, s, code, 5, state.multipoles, state.centers4, loc)
        outs.append(loc)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    C = 36
    a = outs[0][:, :C].double().cpu().numpy()
    b = ref[:, :C].double().cpu().numpy()
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    err = float((np.abs(a - b) / scale).max())
    print(f"[gpu]   max row-relative difference {err:.3e}", flush=True)
    # float32: the two kernels add up to several hundred pair contributions per row in different
    # orders (the per-target kernel by lane, then across lanes; the flat one in 64-pair blocks)
    assert err < (1e-12 if dtype_name == "fp64" else 5e-6)


@pytest.mark.parametrize("learned", [False, True], ids=["plain", "learned_mixed"])
def test_graph_replay_matches_eager(learned, require_gpu):
    """A captured evaluation replays to the eager result, bit for bit, over
    several replays with eager work and new inputs in between."""
    from jaccpot_amd import FastMultipoleMethod

    print(f"[gpu] graph replay learned={learned}", flush=True)
    x, m = ics.plummer(20000, 5)
    x, m = x.astype(np.float32), m.astype(np.float32)
    kw = {}
    if learned:
        from jaccpot_amd.train_order import load

        kw = dict(mixed_precision=True, adaptive_order=True, p_gears=(2, 4, 6), order_model=load(),
                  gear_tolerance=-2.0, order_judge_targets=0)
    fmm = FastMultipoleMethod(precision="fp32", theta=0.6, softening=1e-3, **kw)
    state = fmm.prepare_state(x, m, leaf_size=8, max_order=6)
    xd = torch.as_tensor(x, device="cuda")
    md = torch.as_tensor(m, device="cuda")
    g = fmm.capture_evaluation(state, xd.clone(), md.clone(), return_potential=True)
    rng = np.random.default_rng(0)
    for k in range(3):
        xs = torch.as_tensor(x + rng.normal(scale=1e-4, size=x.shape).astype(np.float32), device="cuda")
        g.positions.copy_(xs)
        acc_g, pot_g = g.replay()
        acc_g, pot_g = acc_g.clone(), pot_g.clone()
        torch.cuda.synchronize()
        # eager evaluation of the same inputs, then compare
        fmm.refresh_prepared_state(state, xs, md)
        acc_e = torch.empty_like(acc_g)
        pot_e = torch.empty_like(pot_g)
        fmm._evaluate_device(state, acc_e, pot_e)
        torch.cuda.synchronize()
        assert torch.equal(acc_g, acc_e) and torch.equal(pot_g, pot_e), k


def _cfmm():
    from oracle import cfmm

    if not cfmm.LIBRARY.exists():
        cfmm.build()
    return cfmm


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_262k_plummer_matches_cpu_restatement(precision, require_gpu):
    """Plummer N=262,144 (p 4, leaf 32, theta 0.6, softening 1e-3): the GPU FMM
    against the C++ restatement in float64 on the same topology (accelerations
    and potentials)."""
    from jaccpot_amd import FastMultipoleMethod

    cfmm = _cfmm()
    print(f"[gpu] 262k plummer vs C++ restatement {precision}", flush=True)
    x, m = ics.plummer(262144, 7)
    if precision == "fp32":
        x, m = x.astype(np.float32), m.astype(np.float32)
    fmm = FastMultipoleMethod(precision=precision, theta=0.6, softening=1e-3)
    state = fmm.prepare_state(x, m, leaf_size=32, max_order=4)
    acc, pot = fmm.evaluate_prepared_state(state, return_potential=True)
    torch.cuda.synchronize()
    ref, rpot = cfmm.evaluate(state.topology, x.astype(np.float64), m.astype(np.float64), p=4, eps=1e-3,
                              return_potential=True)
    ra = np.linalg.norm(acc - ref) / np.linalg.norm(ref)
    rp = np.linalg.norm(pot - rpot) / np.linalg.norm(rpot)
    print(f"[gpu]   rel-L2 acc {ra:.3e} pot {rp:.3e}", flush=True)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert ra < tol and rp < tol


def test_benchmark_workload_matches_cpu_restatement(require_gpu):
    """The benchmark's learned+mixed configuration at full size -- Plummer
    N=4,194,304 (seed 2), float32, leaf 8, theta 0.86, per-pair gears (3, 6, 8)
    from the shipped network at tolerance -1.5, float64 top levels -- against
    the C++ restatement in float64 on the same topology with the same per-pair
    orders: relative L2 difference of the accelerations below 1e-5."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd.train_order import load, weights_path

    cfmm = _cfmm()
    print("[gpu] benchmark workload (4M learned+mixed) vs C++ restatement", flush=True)
    x, m = ics.plummer(4_194_304, 2)
    x32, m32 = x.astype(np.float32), m.astype(np.float32)
    fmm = FastMultipoleMethod(precision="fp32", theta=0.86, softening=1e-3, mixed_precision=True,
                              adaptive_order=True, p_gears=(3, 6, 8), order_model=load(weights_path((3, 6, 8))),
                              gear_tolerance=-1.5)
    state = fmm.prepare_state(x32, m32, leaf_size=8, max_order=8)
    acc = fmm.evaluate_prepared_state(state)
    torch.cuda.synchronize()
    orders = state.extra["gear_of_pair"]
    counts = {int(g): int(np.sum(orders == g)) for g in (3, 6, 8)}
    print(f"[gpu]   GPU done; far pairs {orders.size}, gears {counts}; running the C++ restatement", flush=True)
    assert sum(counts.values()) == orders.size and min(counts.values()) > 0
    ref = cfmm.evaluate(state.topology, x32.astype(np.float64), m32.astype(np.float64), p=8, eps=1e-3,
                        orders=orders)
    rel = np.linalg.norm(acc - ref) / np.linalg.norm(ref)
    per = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"[gpu]   rel-L2 {rel:.3e}; per particle median {np.median(per):.3e} p99 {np.quantile(per, 0.99):.3e}",
          flush=True)
    assert rel < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_split_near_lists_match_whole_leaves(precision, require_gpu):
    """P2P through work items (long near lists cut into slot ranges, partial
    sums added in item order) equals the one-workgroup-per-leaf P2P: to
    rounding in float64, to float32 summation-order noise in float32; with
    potentials, with and without the far field."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd import _native as nat
    from jaccpot_amd.nearsplit import build_p2p_items

    print(f"[gpu] split near lists {precision}", flush=True)
    x, m = ics.plummer(20000, 13)
    if precision == "fp32":
        x, m = x.astype(np.float32), m.astype(np.float32)
    fmm = FastMultipoleMethod(precision=precision, theta=0.7, softening=1e-3, split_near=False)
    state = fmm.prepare_state(x, m, leaf_size=8, max_order=4)
    assert "p2p_items" not in state.extra or state.extra["p2p_items"] is None
    whole, whole_pot = fmm.evaluate_prepared_state(state, return_potential=True)
    near_whole = torch.empty((x.shape[0], 3), dtype=state.dtype, device="cuda")
    fmm._evaluate_device(state, near_whole, None, flags=nat.EVAL_NEAR | nat.EVAL_COMPENSATED)
    items = build_p2p_items(state.topology.near_off, 8, state.dtype, "cuda", chunk_slots=96)
    print(f"[gpu]   {items.num_items} items, {items.num_split} split leaves", flush=True)
    assert items.num_split > 100
    state.extra["p2p_items"] = items
    split, split_pot = fmm.evaluate_prepared_state(state, return_potential=True)
    near_split = torch.empty_like(near_whole)
    fmm._evaluate_device(state, near_split, None, flags=nat.EVAL_NEAR | nat.EVAL_COMPENSATED)
    torch.cuda.synchronize()
    tol = 1e-13 if precision == "fp64" else 2e-6
    for a, b in ((split, whole), (split_pot, whole_pot), (near_split.cpu().numpy(), near_whole.cpu().numpy())):
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < tol, rel


def test_reuse_topology_keeps_tree_while_codes_are_unchanged(require_gpu):
    """reuse_topology / rebuild_every: the tree and lists are kept while the
    sorted Morton codes are unchanged and the budget allows; the reused state
    evaluates like the prepared one on the same topology; changed codes or an
    exhausted budget rebuild."""
    from jaccpot_amd import FastMultipoleMethod

    print("[gpu] reuse topology", flush=True)
    x, m = ics.plummer(20000, 19)
    fmm = FastMultipoleMethod(precision="fp64", theta=0.6, softening=1e-3, reuse_topology=True, rebuild_every=3)
    s0 = fmm.prepare_state(x, m, leaf_size=16, max_order=4)
    a0 = fmm.evaluate_prepared_state(s0)
    # a tiny displacement that leaves every Morton code unchanged
    x1 = x + 1e-13
    s1 = fmm.prepare_state(x1, m, leaf_size=16, max_order=4)
    assert fmm._reuse["count"] == 1 and s1.far_src is s0.far_src and s1 is not s0
    a1 = fmm.evaluate_prepared_state(s1)
    fresh = FastMultipoleMethod(precision="fp64", theta=0.6, softening=1e-3)
    f0 = fresh.prepare_state(x, m, leaf_size=16, max_order=4)
    fresh.refresh_prepared_state(f0, torch.as_tensor(x1, device="cuda"), torch.as_tensor(m, device="cuda"))
    ref1 = fresh.evaluate_prepared_state(f0)
    assert np.array_equal(a1, ref1)
    assert np.array_equal(fmm.evaluate_prepared_state(s0), a0)  # the earlier state is untouched
    s2 = fmm.prepare_state(x1, m, leaf_size=16, max_order=4)
    assert fmm._reuse["count"] == 2 and s2.far_src is s0.far_src
    s3 = fmm.prepare_state(x1, m, leaf_size=16, max_order=4)  # budget of rebuild_every - 1 = 2 reuses spent
    assert fmm._reuse["count"] == 0 and s3.far_src is not s0.far_src
    x4 = x1.copy()
    x4[:100] *= 1.5  # moves particles to other Morton cells
    s4 = fmm.prepare_state(x4, m, leaf_size=16, max_order=4)
    assert fmm._reuse["count"] == 0 and s4.far_src is not s3.far_src


def test_reuse_prepared_state_by_identity(require_gpu):
    from jaccpot_amd import FastMultipoleMethod

    x, m = ics.uniform(3000, 4)
    fmm = FastMultipoleMethod(precision="fp64", theta=0.6, softening=1e-3)
    a = fmm.compute_accelerations(x, m, leaf_size=16, max_order=4)
    st = fmm._last_inputs[3]
    b = fmm.compute_accelerations(x, m, leaf_size=16, max_order=4, reuse_prepared_state=True)
    assert fmm._last_inputs[3] is st and np.array_equal(a, b)
    c = fmm.compute_accelerations(x.copy(), m, leaf_size=16, max_order=4, reuse_prepared_state=True)
    assert fmm._last_inputs[3] is not st and np.array_equal(a, c)


@pytest.mark.parametrize("learned", [False, True], ids=["plain", "learned_mixed"])
def test_saved_state_evaluates_identically(learned, tmp_path, require_gpu):
    """save_prepared_state / load_prepared_state: the loaded state evaluates bit
    for bit like the saved one, and after a refresh with new positions too."""
    from jaccpot_amd import FastMultipoleMethod

    print(f"[gpu] save/load prepared state learned={learned}", flush=True)
    x, m = ics.plummer(20000, 21)
    x, m = x.astype(np.float32), m.astype(np.float32)
    kw, leaf, p = {}, 16, 4
    if learned:
        from jaccpot_amd.train_order import load, weights_path

        kw = dict(mixed_precision=True, adaptive_order=True, p_gears=(3, 6, 8),
                  order_model=load(weights_path((3, 6, 8))), gear_tolerance=-1.5, order_judge_targets=0)
        leaf, p = 8, 8
    fmm = FastMultipoleMethod(precision="fp32", theta=0.8, softening=1e-3, **kw)
    st = fmm.prepare_state(x, m, leaf_size=leaf, max_order=p)
    a, pa = fmm.evaluate_prepared_state(st, return_potential=True)
    path = tmp_path / "state.npz"
    fmm.save_prepared_state(st, path)
    other = FastMultipoleMethod(precision="fp32", softening=0.5)  # settings come from the file
    st2 = other.load_prepared_state(path)
    b, pb = other.evaluate_prepared_state(st2, return_potential=True)
    assert np.array_equal(a, b) and np.array_equal(pa, pb)
    x2 = torch.as_tensor(x * np.float32(1.001), device="cuda")
    md = torch.as_tensor(m, device="cuda")
    fmm.refresh_prepared_state(st, x2, md)
    other.refresh_prepared_state(st2, x2, md)
    assert np.array_equal(fmm.evaluate_prepared_state(st), other.evaluate_prepared_state(st2))


@pytest.mark.parametrize("G", [1.0, 4.3e-3], ids=["G1", "G4.3e-3"])
@pytest.mark.parametrize("learned", [False, True], ids=["plain", "learned_mixed"])
def test_overlapped_step_matches_sequential(learned, G, require_gpu):
    """step_device (P2P on a second stream during the upward/downward passes,
    L2P added afterwards) equals its one-stream order and refresh_prepared_state
    + evaluation bit for bit, also for G != 1 (where near + far must be rounded
    in the same order on every path)."""
    from jaccpot_amd import FastMultipoleMethod

    print(f"[gpu] overlapped step learned={learned} G={G}", flush=True)
    x, m = ics.plummer(50000, 23)
    x, m = x.astype(np.float32), m.astype(np.float32)
    kw, leaf, p = {}, 16, 4
    if learned:
        from jaccpot_amd.train_order import load, weights_path

        kw = dict(mixed_precision=True, adaptive_order=True, p_gears=(3, 6, 8),
                  order_model=load(weights_path((3, 6, 8))), gear_tolerance=-1.5, order_judge_targets=0)
        leaf, p = 8, 8
    fmm = FastMultipoleMethod(precision="fp32", theta=0.8, softening=1e-3, G=G, **kw)
    st = fmm.prepare_state(x, m, leaf_size=leaf, max_order=p)
    xd = torch.as_tensor(x * np.float32(1.0005), device="cuda")
    md = torch.as_tensor(m, device="cuda")
    n = x.shape[0]
    acc_o = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    pot_o = torch.empty((n,), dtype=torch.float32, device="cuda")
    fmm.step_device(st, xd, md, acc_o, pot_o)
    acc_q = torch.empty_like(acc_o)
    pot_q = torch.empty_like(pot_o)
    fmm.step_device(st, xd, md, acc_q, pot_q, overlap=False)
    acc_s = torch.empty_like(acc_o)
    pot_s = torch.empty_like(pot_o)
    fmm.refresh_prepared_state(st, xd, md)
    fmm._evaluate_device(st, acc_s, pot_s)
    torch.cuda.synchronize()
    assert torch.equal(acc_o, acc_q) and torch.equal(pot_o, pot_q)
    assert torch.equal(acc_o, acc_s) and torch.equal(pot_o, pot_s)


def test_graph_replay_after_potential_step(require_gpu):
    """A graph captured without potentials still replays correctly after an
    eager step that asked for them (the deferred far-field rows it captured
    must stay the state's rows)."""
    from jaccpot_amd import FastMultipoleMethod

    print("[gpu] graph replay after a potential step", flush=True)
    x, m = ics.plummer(30000, 6)
    x, m = x.astype(np.float32), m.astype(np.float32)
    fmm = FastMultipoleMethod(precision="fp32", theta=0.7, softening=1e-3)
    st = fmm.prepare_state(x, m, leaf_size=16, max_order=4)
    xd = torch.as_tensor(x, device="cuda")
    md = torch.as_tensor(m, device="cuda")
    g = fmm.capture_evaluation(st, xd.clone(), md.clone(), return_potential=False)
    n = x.shape[0]
    acc_p = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    pot_p = torch.empty((n,), dtype=torch.float32, device="cuda")
    fmm.step_device(st, xd, md, acc_p, pot_p)
    junk = [torch.full((n, 4), 7.0, device="cuda") for _ in range(4)]  # reuse freed blocks, if any
    for k in range(2):
        acc_g = g.replay().clone()
        torch.cuda.synchronize()
        assert torch.equal(acc_g, acc_p), k
    del junk


def test_reuse_prepared_state_ignores_interleaved_prepare(require_gpu):
    """compute_accelerations(reuse_prepared_state=True) reuses the state it
    prepared for the same objects, not whatever prepare_state built last."""
    from jaccpot_amd import FastMultipoleMethod

    x, m = ics.uniform(3000, 7)
    y, m2 = ics.plummer(2500, 8)
    fmm = FastMultipoleMethod(precision="fp64", theta=0.6, softening=1e-3)
    a = fmm.compute_accelerations(x, m, leaf_size=16, max_order=4)
    fmm.prepare_state(y, m2, leaf_size=16, max_order=4)
    b = fmm.compute_accelerations(x, m, leaf_size=16, max_order=4, reuse_prepared_state=True)
    assert b.shape == a.shape and np.array_equal(a, b)


def test_instances_share_the_p2p_stream(require_gpu):
    """Every instance forks the overlapped step's P2P onto one stream per
    device, whatever torch's stream pool has handed out in between: a fresh
    pool stream per instance can share the current stream's hardware queue
    (then the P2P runs after the tree passes instead of beside them)."""
    from jaccpot_amd import FastMultipoleMethod

    print("[gpu] shared P2P stream", flush=True)
    dev = torch.device("cuda", torch.cuda.current_device())
    first = FastMultipoleMethod(precision="fp32")._side_stream(dev)
    pool = [torch.cuda.Stream(device=dev) for _ in range(8)]  # advance the pool
    again = FastMultipoleMethod(precision="fp32")._side_stream(dev)
    assert again is first and all(s is not first for s in pool)
    assert first != torch.cuda.current_stream(dev)
