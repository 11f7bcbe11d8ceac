"""Benchmark of the fixed-topology FMM evaluation path on one MI355X.

One step = one pass of the hot path over the workload at its frozen topology:
gather into tree order, P2M, M2M by level, M2L, L2L by level, L2P and P2P
(the prepare-side numerics plus evaluate_prepared_state).  Inputs are resident
in HBM before the timed region; the topology build (tree + far/near lists)
happens once in prepare and is reported separately.  In the step the P2P
runs on a second stream while the upward and downward passes run
(``FastMultipoleMethod.step_device``; ``--sequential`` puts every launch on
one stream).

Each configuration is timed over K eager steps with HIP events per stage;
the P2P roofline uses the P2P launched alone after the timed steps.
``--graph`` also times K replays of the same step captured once into a HIP
graph and reports those as ``ms_per_step``.  Prints ONE JSON line.  Default workload: Plummer sphere
N=4,194,304 (a=1, truncated at 10a, masses 1/N, seed 2), float32, softening
1e-3; ``value`` is the learned+mixed FMM (leaf 8, theta 0.86, M2L orders
3/6/8 chosen per pair by the order network, float64 top-level M2M/L2L), and
the plain fixed-order FMM at p=6, leaf 64, theta=0.6 and the learned FMM
with gears 2/4/6 (leaf 8, theta 0.7) are timed beside it.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "particles/s at N=4M, median force err ≤1e-4: learned+mixed vs plain GPU FMM"

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    # name: (generator, N, seed, p, leaf, theta, softening, precision)
    "plummer4m_p6": ("plummer", 4_194_304, 2, 6, 64, 0.6, 1e-3, "fp32"),
    "plummer1m_p4": ("plummer", 1_048_576, 1, 4, 32, 0.6, 1e-3, "fp32"),
    "uniform4k_p4_fp64": ("uniform", 4096, 0, 4, 16, 0.6, 0.0, "fp64"),
}


# Learned+mixed configuration per workload: the per-pair order model (gears
# 2/4/6) chooses M2L orders, the top three levels' M2M/L2L run in float64;
# leaf size and opening angle are chosen for the lowest time at median force
# error <= 1e-4 against the FP64 direct sum.
LEARNED = {
    "plummer4m_p6": {"leaf": 8, "theta": 0.86, "gears": (3, 6, 8), "gear_tolerance": -1.5},
    "plummer1m_p4": {"leaf": 8, "theta": 0.85, "gears": (3, 6, 8), "gear_tolerance": -1.5},
}
# Reported beside the value: the learned FMM with the gears (2, 4, 6) of the
# order-model training configuration, at its own fastest setting.
LEARNED_246 = {"leaf": 8, "theta": 0.7, "gears": (2, 4, 6), "gear_tolerance": -1.5}

# The plain fixed-order float32 FMM (no gears, no float64 levels) at the order,
# leaf size and opening angle with the lowest step time at median force error
# <= 1e-4 on the same targets (sweep over p 6..8, leaf 8 and 16, theta
# 0.75..0.92, overlapped step): the like-for-like baseline of the learned+mixed
# configuration.
PLAIN_TUNED = {
    "plummer4m_p6": {"p": 8, "leaf": 16, "theta": 0.9},
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(x, m, p: int, leaf: int, theta: float, eps: float, runs: int = 5):
    """The plain fixed-order FMM timed on the host: oracle/cfmm.cpp (the C++/OpenMP
    restatement of the reference's pipeline), float32 arithmetic, every OpenMP
    thread, evaluations (gather, P2M, M2M, M2L, L2L, L2P + P2P) on the native
    host topology, which is built first and reported separately (as on the GPU).
    One warm-up evaluation, then the median of ``runs`` timed ones."""
    from jaccpot_amd.topology import build_topology
    from oracle import cfmm

    topo = build_topology(x, m, leaf_size=leaf, theta=theta)
    threads = int(cfmm.load().cfmm_max_threads())
    cfmm.evaluate(topo, x, m, p=p, eps=eps, precision="fp32")  # warm-up
    times, stage_runs = [], []
    for _ in range(runs):
        stages = {}
        t0 = time.perf_counter()
        cfmm.evaluate(topo, x, m, p=p, eps=eps, precision="fp32", timings=stages)
        times.append(time.perf_counter() - t0)
        stage_runs.append(stages)
    k = int(np.argsort(times)[len(times) // 2])
    dt = float(times[k])
    return x.shape[0] / dt, dt, threads, stage_runs[k], topo.build_seconds, times


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plummer4m_p6", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="learned",
                    choices=("learned", "plain", "learned-only", "grad-only", "mutual", "configs"),
                    help="learned: time the learned+mixed FMM (value) and the plain one beside it; "
                         "learned-only: the learned+mixed FMM alone (profiling); "
                         "grad-only: the differentiable forward + backward alone (profiling; prints its object); "
                         "mutual: the momentum-conserving block-step lane (prints its own line); "
                         "configs: one evaluation rate per benchmark configuration (prints its own line)")
    ap.add_argument("--mutual-leaf", type=int, default=8, help="leaf size of --mode mutual")
    ap.add_argument("--graph", action="store_true",
                    help="also time the step replayed from a captured HIP graph and report that as ms_per_step")
    ap.add_argument("--no-grad", action="store_true", help="skip the differentiable-path timing")
    ap.add_argument("--sequential", action="store_true",
                    help="run every launch of the step on one stream (no P2P overlap; profiling)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="particles of the CPU baseline run (0: the full workload)")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU baseline runs (median; one warm-up first)")
    ap.add_argument("--error-targets", type=int, default=4096)
    ap.add_argument("--m2l-engine", default="valu", choices=("valu", "mfma"),
                    help="where the float32 M2L applies its B_U rotation blocks (VALU or matrix cores)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one torch copy of this many MiB first (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the FMM path here runs on one GPU (no multi-GPU lane); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from jaccpot_amd import FastMultipoleMethod, direct_sum_gravitational_acceleration
    from jaccpot_amd import _native as nat
    from jaccpot_amd import ics
    from jaccpot_amd.perf import m2l_flops

    if args.mode == "grad-only":
        print(json.dumps({"differentiable": measure_gradient(torch, FastMultipoleMethod, ics, args)}), flush=True)
        return 0
    if args.mode == "mutual":
        print(json.dumps(measure_mutual(torch, ics, args)), flush=True)
        return 0
    if args.mode == "configs":
        print(json.dumps(measure_configs(torch, FastMultipoleMethod, ics, args)), flush=True)
        return 0

    if args.calibrate_copy_mib > 0:
        nel = args.calibrate_copy_mib * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mib} MiB read + {args.calibrate_copy_mib} MiB written")
        del src, dst

    gen, N, seed, p, leaf, theta, eps, precision = WORKLOADS[args.workload]
    log(f"workload {args.workload}: {gen} N={N} p={p} leaf={leaf} theta={theta} eps={eps} {precision}")
    x, m = getattr(ics, gen)(N, seed)
    learned_cfg = LEARNED.get(args.workload) if args.mode in ("learned", "learned-only") else None
    if args.mode == "learned-only" and learned_cfg is None:
        raise SystemExit(f"no learned+mixed configuration for workload {args.workload}")

    def measure(fmm, leaf_size, order, tag):
        """Prepare once, then W warmup + K timed steps with per-stage HIP events."""
        t0 = time.perf_counter()
        state = fmm.prepare_state(x, m, leaf_size=leaf_size, max_order=order)
        torch.cuda.synchronize()
        prep_s = time.perf_counter() - t0
        topo = state.topology
        log(f"[{tag}] prepared in {prep_s:.2f}s (topology build {topo.build_seconds:.2f}s): L={topo.num_leaves} "
            f"nodes={topo.num_nodes} depth={topo.max_depth} far pairs={topo.num_far_pairs} "
            f"near={topo.num_near_entries}")
        dev = state.device
        wdt = state.dtype
        xd = torch.as_tensor(x, device=dev).to(wdt)
        md = torch.as_tensor(m, device=dev).to(wdt)
        acc = torch.empty((N, 3), dtype=wdt, device=dev)

        def run_step(ev=None):
            # gather, then P2P on a second stream while upward + downward run, then L2P adds the far field
            fmm.step_device(state, xd, md, acc, None, events=ev, overlap=not args.sequential)

        for _ in range(args.warmup):
            run_step()
        torch.cuda.synchronize()
        events = [{k: [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                   for k in ("upward", "downward", "l2p", "combine", "p2p")} for _ in range(args.steps)]
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for k in range(args.steps):
            run_step(events[k])
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t_start
        ms_per_step = 1e3 * elapsed / args.steps
        stage = {"upward": [], "downward": [], "l2p": [], "combine": [], "p2p": []}
        for ev in events:
            for k in stage:
                stage[k].append(ev[k][0].elapsed_time(ev[k][1]))
        stage_ms = {k: float(np.mean(v)) for k, v in stage.items()}
        log(f"[{tag}] eager {ms_per_step:.3f} ms/step; stages (ms) "
            f"{json.dumps({k: round(v, 3) for k, v in stage_ms.items()})}")
        eager_ms = ms_per_step
        # the dominant kernel alone (P2P launched by itself, HIP events on its stream): inside the
        # overlapped step it shares the GPU with the tree passes, so its step-stage time is not its own
        p2p_alone = []
        for _ in range(max(args.steps, 3)):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fmm._p2p_device(state, acc, None, nat.EVAL_COMPENSATED, accumulate=False)
            e1.record()
            p2p_alone.append((e0, e1))
        torch.cuda.synchronize()
        p2p_alone_ms = float(np.mean([a.elapsed_time(b) for a, b in p2p_alone]))
        log(f"[{tag}] P2P launched alone: {p2p_alone_ms:.3f} ms")
        if args.graph:
            # the same step captured once into a HIP graph and replayed
            graph = fmm.capture_evaluation(state, xd, md)
            for _ in range(args.warmup):
                graph.replay()
            torch.cuda.synchronize()
            t_start = time.perf_counter()
            for _ in range(args.steps):
                graph.replay()
            torch.cuda.synchronize()
            ms_per_step = 1e3 * (time.perf_counter() - t_start) / args.steps
            log(f"[{tag}] graph replay {ms_per_step:.3f} ms/step")
            del graph
        err = None
        if args.error_targets > 0:
            rng = np.random.default_rng(12345)
            tgt = np.sort(rng.choice(N, size=min(args.error_targets, N), replace=False))
            a_full = torch.empty((N, 3), dtype=wdt, device=dev)
            fmm._evaluate_device(state, a_full, None)
            if "ref" not in ref_cache:  # the same targets for every mode
                ref_cache["ref"] = direct_sum_gravitational_acceleration(x, m, softening=eps, target_indices=tgt)
            ref = ref_cache["ref"]
            got = a_full[torch.as_tensor(tgt, device=dev)].to(torch.float64)
            rel = (torch.linalg.norm(got - ref, dim=1) / torch.linalg.norm(ref, dim=1)).cpu().numpy()
            err = {"median": float(np.median(rel)), "p99": float(np.percentile(rel, 99)), "max": float(rel.max()),
                   "targets": int(tgt.size), "target_seed": 12345}
            log(f"[{tag}] force error vs FP64 direct sum: median {err['median']:.3e} p99 {err['p99']:.3e}")
        sched = state.extra.get("gears")
        far_flops = (sched.m2l_flops if sched is not None else m2l_flops(order) * topo.num_far_pairs)
        info = {"ms_per_step": ms_per_step, "eager_ms_per_step": eager_ms, "graph": bool(args.graph),
                "particles_per_s": N / (ms_per_step * 1e-3), "stage_ms": stage_ms, "p2p_alone_ms": p2p_alone_ms,
                "overlap": not args.sequential,
                "force_error": err, "prepare_s": prep_s, "topology_build_s": topo.build_seconds,
                "leaf_size": leaf_size, "far_pairs": topo.num_far_pairs, "m2l_gflop_per_step": far_flops / 1e9,
                "p2p_interactions": int(topo.p2p_interactions())}
        if sched is not None:
            info["gear_counts"] = {str(k): v for k, v in sched.counts.items()}
        if "order_judge" in state.extra:
            info["order_judge"] = state.extra["order_judge"]
        del state, xd, md, acc
        torch.cuda.empty_cache()
        return info

    ref_cache = {}
    plain = plain_tuned = None
    if args.mode != "learned-only":
        plain = measure(FastMultipoleMethod(precision=precision, theta=theta, softening=eps, compensated=True,
                                            m2l_engine=args.m2l_engine), leaf, p, "plain")
        tuned = PLAIN_TUNED.get(args.workload)
        if tuned is not None and args.mode == "learned":
            plain_tuned = measure(FastMultipoleMethod(precision=precision, theta=tuned["theta"], softening=eps,
                                                      compensated=True, m2l_engine=args.m2l_engine),
                                  tuned["leaf"], tuned["p"], "plain_tuned")
            plain_tuned["config"] = dict(tuned)
    learned = learned246 = None
    if learned_cfg is not None:
        from jaccpot_amd.train_order import load, weights_path

        def learned_fmm(cfg):
            path = weights_path(cfg["gears"])
            if not path.exists():
                raise SystemExit(f"order-model weights {path} missing: run python -m jaccpot_amd.train_order "
                                 f"--gears {','.join(map(str, cfg['gears']))}")
            model = load(path)
            return FastMultipoleMethod(precision=precision, theta=cfg["theta"], softening=eps, compensated=True,
                                       mixed_precision=True, adaptive_order=True, p_gears=tuple(model.gears),
                                       order_model=model, gear_tolerance=cfg["gear_tolerance"],
                                       m2l_engine=args.m2l_engine)

        learned = measure(learned_fmm(learned_cfg), learned_cfg["leaf"], max(learned_cfg["gears"]), "learned")
        if args.mode == "learned":
            learned246 = measure(learned_fmm(LEARNED_246), LEARNED_246["leaf"], max(LEARNED_246["gears"]),
                                 "learned246")
    head = learned if learned is not None else plain
    ms_per_step = head["ms_per_step"]
    stage_ms = head["stage_ms"]
    topo_leaf = head["leaf_size"]

    # ---- roofline of the dominant kernel (P2P, FP32 VALU) -------------------
    interactions = head["p2p_interactions"]
    p2p_flops = 20.0 * interactions
    p2p_s = head["p2p_alone_ms"] * 1e-3
    achieved = p2p_flops / p2p_s / 1e12
    traffic, traffic_note = None, None
    tag = "learned" if learned is not None else "plain"
    tfile = os.path.join(ROOT, "profiles", f"p2p_{args.workload}_{tag}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        if tj.get("p2p_source_sha256") == p2p_source_hash():
            traffic = tj.get("hbm_bytes_per_launch")
            traffic_note = (f"{os.path.relpath(tfile, ROOT)}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this "
                            f"P2P source (sha256 {tj['p2p_source_sha256'][:12]}), scaled by a plain copy kernel")
        else:
            traffic_note = f"{os.path.relpath(tfile, ROOT)} was measured on another P2P source: not reported"
    roofline = {"bound": "valu", "kernel": "p2p_kernel", "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS if precision == "fp32" else 78.6, "unit": "TFLOP/s",
                "frac": round(achieved / (PEAK_FP32_TFLOPS if precision == "fp32" else 78.6), 4),
                "traffic": traffic, "traffic_source": traffic_note,
                "interactions_per_launch": int(interactions), "flops_per_interaction": 20,
                # the kernels are built without packed FP32 (two FMAs per lane per instruction), which is half of
                # the published FP32 peak: against the unpacked FP32 VALU rate the same kernel reads
                "frac_of_unpacked_fp32_valu": round(achieved / 78.6, 4) if precision == "fp32" else None,
                "avg_launch_ms": round(head["p2p_alone_ms"], 4),
                "measured": "P2P (p2p_kernel + split-leaf reduction) launched alone after the timed steps, HIP "
                            "events on its stream; inside the overlapped step it runs beside the tree passes"}

    grad = None
    if args.mode == "learned" and not args.no_grad:
        grad = measure_gradient(torch, FastMultipoleMethod, ics, args)

    cpu = None
    if not args.no_cpu_baseline:
        cn = N if args.cpu_sample <= 0 else min(args.cpu_sample, N)
        xc, mc = (x, m) if cn == N else getattr(ics, gen)(cn, seed)
        rate, dt, threads, stages, topo_s, times = cpu_baseline(xc, mc, p, leaf, theta, eps, args.cpu_runs)
        cpu = {"value": rate, "unit": "particles/s", "cores": threads, "kind": "port",
               "sample": f"C++/OpenMP restatement of the plain fixed-order FMM (oracle/cfmm.cpp), float32, "
                         f"{gen} N={cn} p={p} leaf={leaf} theta={theta} softening={eps}: evaluation "
                         f"(gather, P2M, M2M, M2L, L2L, L2P+P2P), median of {len(times)} after one warm-up, "
                         f"{dt:.2f}s on {threads} threads; host topology build {topo_s:.1f}s not included",
               "runs_s": [round(t, 3) for t in times],
               "stage_s": {k: round(v, 3) for k, v in stages.items()},
               "cpu": platform_cpu(), "nproc": os.cpu_count(),
               "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}
        log(f"cpu baseline {rate:.3e} particles/s on N={cn} ({dt:.1f}s, {threads} threads) "
            f"stages {json.dumps({k: round(v, 2) for k, v in stages.items()})}")

    value = N / (ms_per_step * 1e-3)
    if learned is not None:
        mode = (f"learned per-pair M2L order (gears {learned_cfg['gears']}, tol {learned_cfg['gear_tolerance']}) + "
                f"float64 top-3-level M2M/L2L, leaf={learned_cfg['leaf']} theta={learned_cfg['theta']}, "
                f"compensated P2P")
        workload = (f"{gen} N={N} float32, learned+mixed FMM (leaf={learned_cfg['leaf']}, "
                    f"theta={learned_cfg['theta']}, gears {learned_cfg['gears']}) vs plain fixed-order p={p} "
                    f"leaf={leaf} theta={theta}")
    else:
        mode = "plain fixed-order, compensated P2P"
        workload = f"{gen} N={N} p={p} leaf={leaf} theta={theta} softening={eps} {precision}, fixed-order plain FMM"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": workload, "N": N, "seed": seed,
                   "order": max(learned_cfg["gears"]) if learned is not None else p, "leaf_size": topo_leaf,
                   "softening": eps, "parallelism": "single-gpu", "mode": mode, "m2l_engine": args.m2l_engine},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "force_error": head["force_error"],
        "stage_ms": stage_ms,
        "plain": plain,
        "plain_tuned": plain_tuned,
        "learned_mixed": learned,
        "learned_mixed_gears246": dict(learned246, config=LEARNED_246) if learned246 else None,
        "differentiable": grad,
        "speedup_vs_plain": (learned["particles_per_s"] / plain["particles_per_s"]) if learned and plain else None,
        "speedup_vs_plain_tuned": (learned["particles_per_s"] / plain_tuned["particles_per_s"])
        if learned and plain_tuned else None,
        "p99_force_error": {k: (v["force_error"] or {}).get("p99") for k, v in
                            (("plain", plain), ("plain_tuned", plain_tuned), ("learned_mixed", learned)) if v},
    }
    print(json.dumps(out), flush=True)
    return 0


def measure_configs(torch, FastMultipoleMethod, ics, args):
    """One line per benchmark configuration: K evaluations at a frozen topology
    (gather, upward, downward, L2P, P2P; ``step_device``, overlapped), inputs
    resident on the GPU, and the force error on up to 4096 seeded targets
    against the float64 direct sum.  uniform4k: N=4096 U[-1,1]^3, masses
    U[0.5,1.5], float64, p 4, leaf 16, theta 0.6, no softening;
    plummer1m: N=1,048,576 Plummer, float32, p 4, leaf 32, theta 0.6,
    softening 1e-3; galaxy2m: N=2,097,152 (Hernquist halo 80% + exponential
    disc 20%), float32, softening 1e-3, leaf 16, theta 0.7 -- the learned
    gears (2, 4, 6) with float64 top levels (tolerance -2.5) and the plain
    order-6 FMM; blobs1m: the differentiable forward + backward of
    ``measure_gradient``."""
    from jaccpot_amd import direct_sum_gravitational_acceleration

    def run(tag, x, m, fmm, leaf, order, eps):
        t0 = time.perf_counter()
        st = fmm.prepare_state(x, m, leaf_size=leaf, max_order=order)
        torch.cuda.synchronize()
        prep = time.perf_counter() - t0
        n = x.shape[0]
        xd = torch.as_tensor(x, device="cuda").to(st.dtype)
        md = torch.as_tensor(m, device="cuda").to(st.dtype)
        acc = torch.empty((n, 3), dtype=st.dtype, device="cuda")
        for _ in range(max(1, args.warmup)):
            fmm.step_device(st, xd, md, acc, None)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.steps):
            fmm.step_device(st, xd, md, acc, None)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / args.steps
        tgt = np.sort(np.random.default_rng(12345).choice(n, size=min(4096, n), replace=False))
        ref = direct_sum_gravitational_acceleration(x, m, softening=eps, target_indices=tgt)
        got = acc[torch.as_tensor(tgt, device="cuda")].to(torch.float64)
        rel = (torch.linalg.norm(got - ref, dim=1) / torch.linalg.norm(ref, dim=1)).cpu().numpy()
        out = {"ms_per_step": ms, "particles_per_s": n / (ms * 1e-3), "prepare_s": prep,
               "force_error": {"median": float(np.median(rel)), "p99": float(np.percentile(rel, 99)),
                               "targets": int(tgt.size)}}
        sched = st.extra.get("gears")
        if sched is not None:
            out["gear_counts"] = {str(k): v for k, v in sched.counts.items()}
        log(f"[{tag}] {ms:.3f} ms, median err {out['force_error']['median']:.2e}")
        del st, xd, md, acc
        torch.cuda.empty_cache()
        return out

    res = {}
    x, m = ics.uniform(4096, 0)
    res["uniform4k_p4_fp64"] = run("uniform4k", x, m, FastMultipoleMethod(precision="fp64", theta=0.6), 16, 4, 0.0)
    x, m = ics.plummer(1_048_576, 1)
    res["plummer1m_p4_fp32"] = run("plummer1m", x.astype(np.float32), m.astype(np.float32),
                                   FastMultipoleMethod(precision="fp32", theta=0.6, softening=1e-3), 32, 4, 1e-3)
    x, m = ics.galaxy(2_097_152, 3)
    x32, m32 = x.astype(np.float32), m.astype(np.float32)
    res["galaxy2m_learned246_mixed"] = run(
        "galaxy2m learned", x32, m32,
        FastMultipoleMethod(precision="fp32", theta=0.7, softening=1e-3, adaptive_order=True, p_gears=(2, 4, 6),
                            gear_tolerance=-2.5, mixed_precision=True), 16, 6, 1e-3)
    res["galaxy2m_plain_p6"] = run("galaxy2m plain", x32, m32,
                                   FastMultipoleMethod(precision="fp32", theta=0.7, softening=1e-3), 16, 6, 1e-3)
    res["blobs1m_grad_p4_fp32"] = measure_gradient(torch, FastMultipoleMethod, ics, args)
    return {"configs": res, "steps": args.steps, "warmup": args.warmup}


def measure_mutual(torch, ics, args):
    """The momentum-conserving block-step lane (jaccpot_amd.mutual) on a
    Hernquist sphere (a=1, truncated at 100a, N=1,048,576, equal masses, seed
    3), float64, order 4, theta 0.7, softening 1e-3, leaf ``--mutual-leaf``:
    K full traversals (all levels, unit weights) and K fused boundary
    traversals (rungs 0..3 from the acceleration magnitude, the weights of
    sub-step boundary 1 at k_max 3), inputs and outputs resident on the GPU;
    momentum residual and force error (4096 seeded targets against the
    float64 direct sum) of the full traversal."""
    from jaccpot_amd import direct_sum_gravitational_acceleration
    from jaccpot_amd import mutual as mu

    n, theta, order, eps, leaf, k_max = 1_048_576, 0.7, 4, 1e-3, args.mutual_leaf, 3
    x, m = ics.galaxy(n, 3, halo_fraction=1.0)
    t0 = time.perf_counter()
    topo, _ = mu.build_mutual_topology(x, m, theta=theta, order=order, leaf_size=leaf)
    st = mu.build_mutual_state(topo, softening=eps)
    torch.cuda.synchronize()
    prep_s = time.perf_counter() - t0
    xd = torch.as_tensor(x, device="cuda")
    md = torch.as_tensor(m, device="cuda")
    acc = mu.mutual_accelerations(st, xd, md)
    amag = torch.linalg.norm(acc, dim=1)
    rung = torch.clamp(torch.floor(torch.log2(torch.clamp(amag / amag.median(), min=1.0))), 0, k_max).to(torch.int32)
    w = mu.boundary_level_weights(1, k_max, 0.01)

    def timed(fn):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / args.steps

    ms_full = timed(lambda: mu.mutual_accelerations(st, xd, md))
    ms_kick = timed(lambda: mu.mutual_weighted_accelerations(st, xd, md, rung=rung, level_weights=w))
    acc = mu.mutual_accelerations(st, xd, md)
    terms = md[:, None] * acc
    momentum = float(torch.linalg.norm(terms.sum(0)) / torch.linalg.norm(terms.abs().sum(0)))
    tgt = np.sort(np.random.default_rng(12345).choice(n, 4096, replace=False))
    ref = direct_sum_gravitational_acceleration(xd, md, G=1.0, softening=eps, target_indices=tgt)
    a_t = acc[torch.as_tensor(tgt, device="cuda")]
    err = (torch.linalg.norm(a_t - ref, dim=1) / torch.linalg.norm(ref, dim=1)).cpu().numpy()
    interactions = int(st.prepared.topology.p2p_interactions())
    counts = np.bincount(rung.cpu().numpy(), minlength=k_max + 1).tolist()
    log(f"mutual: full {ms_full:.3f} ms, fused boundary {ms_kick:.3f} ms, momentum {momentum:.2e}, "
        f"median err {np.median(err):.2e}")
    return {
        "metric": "mutual-lane traversals at N=1M Hernquist, float64 (momentum-conserving, block-step weights)",
        "value": n / (ms_full * 1e-3), "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_full, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Hernquist a=1 r<=100a N={n} seed 3, theta {theta}, order {order}, leaf {leaf}, "
                               f"softening {eps}, float64", "parallelism": "single-gpu"},
        "ms_fused_boundary": ms_kick, "rung_counts": counts, "prepare_s": prep_s,
        "momentum_residual": momentum,
        "force_error": {"median": float(np.median(err)), "p99": float(np.percentile(err, 99)),
                        "targets": int(tgt.size)},
        "near_interactions": interactions, "far_pairs_canonical": topo.num_far_pairs,
        "near_pairs_canonical": topo.num_near_pairs,
    }


def measure_gradient(torch, FastMultipoleMethod, ics, args):
    """Forward + backward of the differentiable FMM on 64 Gaussian blobs
    (N=1,048,576, sigma 0.05, lognormal masses, seed 4), float32, p=4, leaf 32,
    theta 0.6, softening 1e-3; loss = sum_t m_t phi_t + 0.5 |a_t|^2 over 1024
    seeded targets; gradients with respect to all positions and masses."""
    n = 1_048_576
    x, m = ics.blobs(n, 4)
    fmm = FastMultipoleMethod(precision="fp32", theta=0.6, softening=1e-3)
    t0 = time.perf_counter()
    state = fmm.prepare_state(x.astype(np.float32), m.astype(np.float32), leaf_size=32, max_order=4)
    torch.cuda.synchronize()
    prep = time.perf_counter() - t0
    tgt = torch.as_tensor(np.sort(np.random.default_rng(12345).choice(n, 1024, replace=False)), device="cuda")
    xt = torch.as_tensor(x, dtype=torch.float32, device="cuda").requires_grad_(True)
    mt = torch.as_tensor(m, dtype=torch.float32, device="cuda").requires_grad_(True)

    def step(backward: bool):
        acc, pot = fmm.differentiable_accelerations_and_potential(state, xt, mt)
        loss = (mt[tgt] * pot[tgt]).sum() + 0.5 * (acc[tgt] * acc[tgt]).sum()
        if backward:
            xt.grad = None
            mt.grad = None
            loss.backward()
        return loss

    for _ in range(max(1, args.warmup)):
        step(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(False)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    both = (time.perf_counter() - t0) / args.steps
    log(f"[grad] blobs N={n} p=4 leaf=32: forward {1e3 * fwd:.2f} ms, forward+backward {1e3 * both:.2f} ms")
    return {"workload": "64 Gaussian blobs N=1048576 sigma 0.05, lognormal masses, seed 4, float32, p=4, leaf 32, "
                        "theta 0.6, softening 1e-3; loss over 1024 targets",
            "forward_ms": 1e3 * fwd, "forward_backward_ms": 1e3 * both, "prepare_s": prep,
            "particles_per_s_forward_backward": n / both}


def p2p_source_hash() -> str:
    """sha256 of the P2P kernel source (ties a committed traffic figure to the kernel it measured)."""
    import hashlib

    with open(os.path.join(ROOT, "jaccpot_amd", "csrc", "p2p.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def platform_cpu() -> str:
    """CPU model string with family/model/stepping and socket count from /proc/cpuinfo."""
    info, sockets = {}, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    continue
                k, v = (t.strip() for t in line.split(":", 1))
                if k == "physical id":
                    sockets.add(v)
                elif k in ("model name", "cpu family", "model", "stepping", "cpu MHz") and k not in info:
                    info[k] = v
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    return (f"{name} (family {info.get('cpu family', '?')} model {info.get('model', '?')} stepping "
            f"{info.get('stepping', '?')}, {len(sockets) or '?'} socket(s), {info.get('cpu MHz', '?')} MHz)")


if __name__ == "__main__":
    sys.exit(main())
