"""GPU: per-gear M2L, the order MLP's kernels against the NumPy oracle, and
what the trained network gains over the fixed-order method."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fmm as ofmm
from oracle import harmonics as H
from oracle import ics
from oracle import mlp as omlp
from oracle import topology as OT

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_gear_m2l_matches_oracle(require_gpu):
    """Random per-pair gears (fp64): locals equal the oracle's truncated-multipole M2L."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd import _native as nat
    from jaccpot_amd.learned import build_gear_schedule

    print("[gpu] gear m2l fp64", flush=True)
    x, m = ics.plummer(3000, 21)
    p, leaf, theta = 6, 16, 0.6
    fmm = FastMultipoleMethod(precision="fp64", theta=theta, softening=1e-3)
    state = fmm.prepare_state(x, m, leaf_size=leaf, max_order=p)
    K = state.topology.num_far_pairs
    g = np.random.default_rng(3).choice([2, 4, 6], size=K).astype(np.int8)
    state.extra["gears"] = build_gear_schedule(state, g, (2, 4, 6))
    fmm._downward(state, nat.current_stream(), nat.JFM_F64)
    torch.cuda.synchronize()
    topo = OT.build_topology(x, m, leaf, theta)
    M = ofmm.upward(topo, x[topo.perm], m[topo.perm], p)
    Lc = ofmm.downward(topo, M, p, orders=g.astype(np.int64))
    gl = state.locals.cpu().numpy()[:, : H.sh_size(p)]
    scale = np.maximum(np.abs(Lc).max(axis=1, keepdims=True), 1e-300)
    assert float((np.abs(gl - Lc) / scale).max()) < 1e-12


@pytest.mark.parametrize("gears", [(2, 4, 6), (3, 6, 8)], ids=["g246", "g368"])
def test_pair_labels_match_oracle(gears, require_gpu):
    """GPU labels (float32, order-8 upward on the device) against the float64
    label oracle on the same topology, where the error is above float32 noise."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd.learned import pair_labels
    from oracle import labels as OL

    print("[gpu] pair labels", flush=True)
    x, m = ics.plummer(3000, 23)
    leaf, theta = 16, 0.6
    fmm = FastMultipoleMethod(precision="fp32", theta=theta, softening=1e-3)
    state = fmm.prepare_state(x.astype(np.float32), m.astype(np.float32), leaf_size=leaf, max_order=6)
    got = pair_labels(state, gears).cpu().numpy().astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    m32 = m.astype(np.float32).astype(np.float64)
    topo = OT.build_topology(x32, m32, leaf, theta)
    assert np.array_equal(topo.far_src, state.topology.far_src)
    M8 = ofmm.upward(topo, x32[topo.perm], m32[topo.perm], 8)
    tgt = np.repeat(np.arange(topo.num_nodes), np.diff(topo.far_off))
    src = topo.far_src
    want = OL.pair_labels(M8[src], topo.center[tgt] - topo.center[src], topo.radius[tgt], gears=gears)
    if gears[-1] == 8:
        assert np.abs(got[:, -1] - np.log10(OL.LABEL_FLOOR)).max() < 1e-5
    sel = want > -4.0
    assert sel.mean() > 0.2
    diff = np.abs(got - want)[sel]
    print(f"[gpu]   {sel.sum()} labels above 1e-4: max |diff| {diff.max():.3e} (log10)", flush=True)
    assert diff.max() < 0.02


def test_mlp_kernels_match_oracle(require_gpu):
    from jaccpot_amd.learned import OrderModel

    print("[gpu] mlp fwd/bwd", flush=True)
    model = OrderModel(seed=5)
    rng = np.random.default_rng(1)
    K = 1537
    X = rng.normal(size=(K, 8)).astype(np.float32)
    T = rng.normal(size=(K, 3)).astype(np.float32)
    Xd = torch.as_tensor(X, device="cuda")
    Td = torch.as_tensor(T, device="cuda")
    Y = model.predict(Xd, normalised=True, exact=True).cpu().numpy()
    theta = omlp.bf16_round(model.master.cpu().numpy())
    Yr, _ = omlp.forward(theta, X)
    assert np.abs(Y - Yr).max() <= 2e-5 * max(1.0, np.abs(Yr).max())
    loss, grad = model.loss_and_grad(Xd, Td, per_block=200)
    lr_, gr = omlp.loss_and_grad(theta, X, T)
    assert abs(float(loss) - lr_) <= 1e-5 * lr_
    g = grad.cpu().numpy()
    assert np.linalg.norm(g - gr) <= 1e-4 * np.linalg.norm(gr)
    # determinism: identical gradient bits on a second call
    _, grad2 = model.loss_and_grad(Xd, Td, per_block=200)
    assert torch.equal(grad, grad2)
    # Adam step moves the masters against the gradient
    before = model.master.clone()
    model.adam(grad, lr=1e-3)
    step = (model.master - before).cpu().numpy()
    assert np.all(step * g <= 1e-12)


def test_mfma_forward_matches_oracle(require_gpu):
    """The matrix-core forward (bf16 activations between layers) against the
    float64 oracle with the same bf16 weights, on random and on real features."""
    from jaccpot_amd import _native as nat
    from jaccpot_amd.learned import OrderModel

    print("[gpu] mlp forward on MFMA", flush=True)
    model = OrderModel(seed=9)
    rng = np.random.default_rng(4)
    for K in (1, 31, 32, 33, 4097, 70001):
        X = rng.normal(size=(K, 8)).astype(np.float32)
        Xd = torch.as_tensor(X, device="cuda")
        Y = torch.empty((K, 3), dtype=torch.float32, device="cuda")
        nat.call("jfm_mlp_forward_mfma", nat.current_stream(), K, Xd.data_ptr(), None, None, model.wbf16.data_ptr(),
                 Y.data_ptr())
        theta = omlp.bf16_round(model.master.cpu().numpy())
        Yr, _ = omlp.forward(theta, X)
        err = float(np.abs(Y.cpu().numpy() - Yr).max())
        scale = max(1.0, float(np.abs(Yr).max()))
        print(f"[gpu]   K={K}: max |Y - Y_oracle| {err:.3e} (scale {scale:.2f})", flush=True)
        assert err <= 3e-2 * scale
    # in-kernel normalisation: same result as normalising on the host first
    mean = rng.normal(size=8).astype(np.float32)
    std = rng.uniform(0.5, 2.0, size=8).astype(np.float32)
    model._set_stats(mean, std)
    X = rng.normal(size=(5000, 8)).astype(np.float32) * std + mean
    Xd = torch.as_tensor(X, device="cuda")
    Y1 = torch.empty((5000, 3), dtype=torch.float32, device="cuda")
    Y2 = torch.empty_like(Y1)
    nat.call("jfm_mlp_forward_mfma", nat.current_stream(), 5000, Xd.data_ptr(), model.feature_mean.data_ptr(),
             model.feature_inv_std.data_ptr(), model.wbf16.data_ptr(), Y1.data_ptr())
    Xn = torch.as_tensor(((X - mean) * (1.0 / std)).astype(np.float32), device="cuda")
    nat.call("jfm_mlp_forward_mfma", nat.current_stream(), 5000, Xn.data_ptr(), None, None, model.wbf16.data_ptr(),
             Y2.data_ptr())
    assert torch.allclose(Y1, Y2, atol=1e-5, rtol=0)


def _train_model(seed=0):
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd.learned import OrderModel, pair_features, pair_labels

    x, m = ics.galaxy(131072, 31 + seed)
    fmm = FastMultipoleMethod(precision="fp32", theta=0.6, softening=1e-3)
    st = fmm.prepare_state(x.astype(np.float32), m.astype(np.float32), leaf_size=32, max_order=6)
    X = pair_features(st)
    T = pair_labels(st)
    model = OrderModel(seed=seed)
    hist = model.fit(X, T, steps=800, batch=16384, lr=3e-3, log_every=100)
    return model, hist, X, T


def test_trained_network_gains_over_fixed_order(require_gpu):
    """Training lowers the loss; at equal accuracy the learned gears cut M2L work."""
    from jaccpot_amd import FastMultipoleMethod, direct_sum_gravitational_acceleration
    from jaccpot_amd.perf import m2l_flops

    print("[gpu] training order model", flush=True)
    model, hist, X, T = _train_model()
    print(f"[gpu]   loss {hist[0][1]:.4f} -> {hist[-1][1]:.4f}", flush=True)
    assert hist[-1][1] < 0.25 * hist[0][1]
    # held-out prediction quality on the training configuration
    Y = model.predict(X)
    resid = (Y - T).abs().median().item()
    print(f"[gpu]   median |log10 residual| {resid:.3f}", flush=True)
    assert resid < 0.5, resid

    x, m = ics.galaxy(65536, 77)
    x32, m32 = x.astype(np.float32), m.astype(np.float32)
    tgt = np.random.default_rng(12345).choice(x.shape[0], 2048, replace=False)
    ref = direct_sum_gravitational_acceleration(x32.astype(np.float64), m32.astype(np.float64), softening=1e-3,
                                                target_indices=tgt).cpu().numpy()

    def err(acc):
        return float(np.median(np.linalg.norm(acc[tgt] - ref, axis=1) / np.linalg.norm(ref, axis=1)))

    plain = FastMultipoleMethod(precision="fp32", theta=0.6, softening=1e-3)
    st = plain.prepare_state(x32, m32, leaf_size=32, max_order=6)
    e_plain = err(plain.evaluate_prepared_state(st))
    K = st.topology.num_far_pairs
    plain_flops = m2l_flops(6) * K
    best = None
    for tol in (-3.0, -2.5, -2.25, -2.0, -1.75, -1.5, -1.25, -1.0):
        learned = FastMultipoleMethod(precision="fp32", theta=0.6, softening=1e-3, adaptive_order=True,
                                      p_gears=(2, 4, 6), order_model=model, gear_tolerance=tol,
                                      order_judge_targets=0)
        stl = learned.prepare_state(x32, m32, leaf_size=32, max_order=6)
        e = err(learned.evaluate_prepared_state(stl))
        sched = stl.extra["gears"]
        print(f"[gpu]   tol {tol}: plain err {e_plain:.3e}; learned err {e:.3e}; gears {sched.counts}; "
              f"M2L flops {sched.m2l_flops / plain_flops:.3f} of fixed order", flush=True)
        if e <= 1.5 * e_plain + 2e-6 and (best is None or sched.m2l_flops < best[1].m2l_flops):
            best = (e, sched)
            best_orders = np.asarray(stl.extra["gear_of_pair"]).copy()
    assert best is not None
    e_learned, sched = best
    assert e_learned <= 1.5 * e_plain + 2e-6
    assert sched.m2l_flops <= 0.85 * plain_flops
    # the network beats an order rule without it (all pairs at the middle gear) on accuracy
    g4 = np.full(K, 4, np.int8)
    from jaccpot_amd.learned import build_gear_schedule

    st.extra["gears"] = build_gear_schedule(st, g4, (2, 4, 6))
    from jaccpot_amd import _native as nat

    plain._downward(st, nat.current_stream(), nat.JFM_F32)
    e_g4 = err(plain.evaluate_prepared_state(st))
    print(f"[gpu]   all pairs at gear 4: err {e_g4:.3e}", flush=True)
    assert e_learned < e_g4
    # the same gear counts (the same M2L flops) dealt to random pairs: the network's
    # choice of which pairs get the low gears is what keeps the error down
    chosen = best_orders
    shuffled = np.random.default_rng(3).permutation(chosen).astype(np.int8)
    st.extra["gears"] = build_gear_schedule(st, shuffled, (2, 4, 6))
    plain._downward(st, nat.current_stream(), nat.JFM_F32)
    e_shuf = err(plain.evaluate_prepared_state(st))
    print(f"[gpu]   learned gear counts dealt at random: err {e_shuf:.3e} (learned {e_learned:.3e})", flush=True)
    assert e_shuf > 2.0 * e_learned


def test_shipped_gear368_model_cuts_order8_work(require_gpu):
    """The shipped (3, 6, 8) network (top gear = the labels' order-8 reference,
    budget extrapolated from the two lower gears) keeps the force error of the
    all-order-8 method within 1.5x while sending a share of the far pairs to
    orders 6 and 3, on a configuration it was not trained on (Plummer)."""
    from jaccpot_amd import FastMultipoleMethod, direct_sum_gravitational_acceleration
    from jaccpot_amd.learned import build_gear_schedule
    from jaccpot_amd.perf import m2l_flops
    from jaccpot_amd.train_order import load, weights_path

    print("[gpu] shipped (3,6,8) model", flush=True)
    model = load(weights_path((3, 6, 8)))
    x, m = ics.plummer(131072, 8)
    x32, m32 = x.astype(np.float32), m.astype(np.float32)
    tgt = np.random.default_rng(12345).choice(x.shape[0], 2048, replace=False)
    ref = direct_sum_gravitational_acceleration(x32.astype(np.float64), m32.astype(np.float64), softening=1e-3,
                                                target_indices=tgt).cpu().numpy()

    def err(acc):
        return float(np.median(np.linalg.norm(acc[tgt] - ref, axis=1) / np.linalg.norm(ref, axis=1)))

    plain = FastMultipoleMethod(precision="fp32", theta=0.85, softening=1e-3)
    st = plain.prepare_state(x32, m32, leaf_size=8, max_order=8)
    e8 = err(plain.evaluate_prepared_state(st))
    K = st.topology.num_far_pairs
    learned = FastMultipoleMethod(precision="fp32", theta=0.85, softening=1e-3, adaptive_order=True,
                                  p_gears=(3, 6, 8), order_model=model, gear_tolerance=-1.5, order_judge_targets=0)
    stl = learned.prepare_state(x32, m32, leaf_size=8, max_order=8)
    el = err(learned.evaluate_prepared_state(stl))
    sched = stl.extra["gears"]
    frac = sched.m2l_flops / (m2l_flops(8) * K)
    print(f"[gpu]   order 8: err {e8:.3e}; learned: err {el:.3e}, gears {sched.counts}, M2L flops {frac:.3f}",
          flush=True)
    assert el <= 1.5 * e8 + 2e-6
    assert frac <= 0.9
    # without the extrapolated budget (every pair at gear 6) the error is far larger
    st.extra["gears"] = build_gear_schedule(st, np.full(K, 6, np.int8), (3, 6, 8))
    from jaccpot_amd import _native as nat

    plain._downward(st, nat.current_stream(), nat.JFM_F32)
    e6 = err(plain.evaluate_prepared_state(st))
    print(f"[gpu]   all pairs at order 6: err {e6:.3e}", flush=True)
    assert el < e6


def test_leaf_budget_never_raises_a_gear(require_gpu):
    """gear_budget="leaf": a target's budget is the smallest, over its leaves,
    of the largest top-gear proxy on the root-to-leaf path, which is never
    below the target's own largest proxy (the "node" budget), so every pair's
    gear is at most its node-budget gear; the budget arrays checked against
    a NumPy pass over the same node maxima."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd import _native as nat
    from jaccpot_amd.learned import pair_features
    from jaccpot_amd.train_order import load, weights_path

    print("[gpu] leaf gear budget vs node budget", flush=True)
    x, m = ics.plummer(65_536, 17)
    x32, m32 = x.astype(np.float32), m.astype(np.float32)
    model = load(weights_path((3, 6, 8)))
    fmm = FastMultipoleMethod(precision="fp32", theta=0.85, softening=1e-3)
    st = fmm.prepare_state(x32, m32, leaf_size=8, max_order=8)
    X = pair_features(st)
    g_node = model.select_gears(X, -1.5, st, budget="node").cpu().numpy()
    g_leaf = model.select_gears(X, -1.5, st, budget="leaf").cpu().numpy()
    print(f"[gpu]   node {np.bincount(g_node, minlength=9)[[3, 6, 8]]} leaf {np.bincount(g_leaf, minlength=9)[[3, 6, 8]]}",
          flush=True)
    assert np.all(g_leaf <= g_node) and np.any(g_leaf < g_node)
    # the budget pass against NumPy on the same node maxima
    topo = st.topology
    Y = model.predict(X)
    g3 = np.asarray((3, 6, 8), np.int32)
    b = torch.empty(topo.num_nodes, dtype=torch.float32, device="cuda")
    nat.call("jfm_gear_node_max", nat.current_stream(), topo.num_nodes, st.far_off.data_ptr(), st.far_src.data_ptr(),
             Y.data_ptr(), st.centers4.data_ptr(), nat.ptr(g3), b.data_ptr())
    node_max = b.cpu().numpy().astype(np.float64)
    path = node_max.copy()
    for u in np.argsort(topo.depth, kind="stable"):  # root first
        if topo.parent[u] >= 0:
            path[u] = max(path[u], path[topo.parent[u]])
    want = path.copy()
    for u in np.argsort(-topo.depth[: topo.num_internal], kind="stable"):  # internal nodes, deepest first
        want[u] = min(want[topo.left[u]], want[topo.right[u]])
    off = topo.l2l_off
    for lev in range(len(off) - 1):
        a, e = int(off[lev]), int(off[lev + 1])
        if e > a:
            nat.call("jfm_level_path_max", nat.current_stream(), st.l2l_nodes.data_ptr() + 4 * a, e - a,
                     st.parent.data_ptr(), b.data_ptr())
    off = topo.m2m_off
    for lev in range(len(off) - 1):
        a, e = int(off[lev]), int(off[lev + 1])
        if e > a:
            nat.call("jfm_level_subtree_min", nat.current_stream(), st.m2m_nodes.data_ptr() + 4 * a, e - a,
                     st.left.data_ptr(), st.right.data_ptr(), b.data_ptr())
    got = b.cpu().numpy().astype(np.float64)
    finite = np.isfinite(want)
    assert np.array_equal(np.isfinite(got), finite)
    assert np.array_equal(got[finite], want[finite])
    assert np.all(got[finite] >= node_max[finite])


def test_gear_schedule_matches_host_split(require_gpu):
    """The device gear split (radix sort by gear, CSR per gear, targets, flat
    lists and their spanning targets) equals a NumPy split of the far CSR."""
    from jaccpot_amd import FastMultipoleMethod
    from jaccpot_amd.flatlist import BLOCK, build_flat_list
    from jaccpot_amd.learned import build_gear_schedule

    print("[gpu] gear schedule split", flush=True)
    x, m = ics.plummer(6000, 9)
    fmm = FastMultipoleMethod(precision="fp32", theta=0.5, softening=1e-3)
    state = fmm.prepare_state(x.astype(np.float32), m.astype(np.float32), leaf_size=4, max_order=6)
    topo = state.topology
    K, nodes = topo.num_far_pairs, topo.num_nodes
    g = np.random.default_rng(5).choice([2, 4, 6], size=K, p=[0.6, 0.3, 0.1]).astype(np.int8)
    sched = build_gear_schedule(state, g, (2, 4, 6))
    tgt = np.repeat(np.arange(nodes), np.diff(topo.far_off))

    def host_spans(t):
        if t.size == 0:
            return set()
        starts = np.concatenate([[0], np.nonzero(np.diff(t))[0] + 1])
        ends = np.concatenate([starts[1:], [t.size]]) - 1
        w0, w1 = starts // BLOCK, ends // BLOCK
        sp = w1 > w0
        return set(zip(t[starts[sp]].tolist(), w0[sp].tolist(), w1[sp].tolist()))

    for gi, gear in enumerate((2, 4, 6)):
        sel = g == gear
        t, s = tgt[sel], topo.far_src[sel]
        cnt = np.bincount(t, minlength=nodes)
        off = np.concatenate([[0], np.cumsum(cnt)])
        tl = np.nonzero(cnt)[0]
        assert sched.counts[gear] == int(sel.sum())
        assert sched.num_targets[gi] == tl.size
        assert np.array_equal(sched.offsets[gi].cpu().numpy(), off)
        assert np.array_equal(sched.sources[gi].cpu().numpy()[: t.size], s)
        assert np.array_equal(sched.targets[gi].cpu().numpy()[: tl.size], tl)
        fl = build_flat_list(t, s, gear, torch.float32, "cuda")
        assert np.array_equal(fl.pair_tgt.cpu().numpy()[: t.size], t)
        assert np.array_equal(fl.pair_src.cpu().numpy()[: t.size], s)
        got = set(zip(*(a.cpu().numpy()[: fl.num_span].tolist() for a in (fl.span_tgt, fl.span_w0, fl.span_w1))))
        assert got == host_spans(t) and fl.num_span == len(got)
        print(f"[gpu]   gear {gear}: {t.size} pairs, {tl.size} targets, {fl.num_span} spanning targets", flush=True)
