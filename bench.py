"""Benchmark of the fixed-topology FMM evaluation path on one MI355X.

One step = one pass of the hot path over the workload at its frozen topology:
gather into tree order, P2M, M2M by level, M2L, L2L by level, L2P and P2P
(the prepare-side numerics plus evaluate_prepared_state).  Inputs are resident
in HBM before the timed region; the host topology build (tree + far/near
lists) happens once in prepare and is reported separately.

Prints ONE JSON line (rank 0).  Default workload: Plummer sphere N=4,194,304
(a=1, truncated at 10a, masses 1/N, seed 2), float32, p=6, leaf 64,
theta=0.6, softening 1e-3.
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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline_sample(gen: str, p: int, leaf: int, theta: float, eps: float, n_sample: int):
    """The float64 oracle FMM (NumPy restatement, single thread) on a bounded sample."""
    from oracle import fmm as ofmm
    from oracle import ics as oics

    x, m = getattr(oics, gen)(n_sample, 12)
    t0 = time.perf_counter()
    ofmm.fmm_accelerations(x, m, leaf_size=leaf, p=p, theta=theta, eps=eps)
    dt = time.perf_counter() - t0
    return n_sample / dt, dt


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plummer4m_p6", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=16384)
    ap.add_argument("--error-targets", type=int, default=4096)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one torch copy of this many MiB first (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md for how it would "
              "shard); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from jaccpot_amd import FastMultipoleMethod, direct_sum_gravitational_acceleration
    from jaccpot_amd import _native as nat
    from jaccpot_amd import ics
    from jaccpot_amd.perf import m2l_flops

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
    fmm = FastMultipoleMethod(precision=precision, theta=theta, softening=eps, compensated=True)
    t0 = time.perf_counter()
    state = fmm.prepare_state(x, m, leaf_size=leaf, max_order=p)
    torch.cuda.synchronize()
    prep_s = time.perf_counter() - t0
    topo = state.topology
    log(f"prepared in {prep_s:.2f}s (host topology {topo.build_seconds:.2f}s): L={topo.num_leaves} "
        f"nodes={topo.num_nodes} depth={topo.max_depth} far pairs={topo.num_far_pairs} near={topo.num_near_entries}")

    dev = state.device
    wdt = state.dtype
    xd = torch.as_tensor(x, device=dev).to(wdt)
    md = torch.as_tensor(m, device=dev).to(wdt)
    acc = torch.empty((N, 3), dtype=wdt, device=dev)
    code = nat.dtype_code(wdt)
    flags = nat.EVAL_COMPENSATED

    def run_step(ev=None):
        s = nat.current_stream()
        if ev is not None:
            ev[0].record()
        nat.call("jfm_gather_sorted", s, code, code, xd.data_ptr(), md.data_ptr(), state.perm.data_ptr(), N,
                 state.pos4.data_ptr())
        fmm._upward(state, s, code)
        if ev is not None:
            ev[1].record()
        fmm._downward(state, s, code)
        if ev is not None:
            ev[2].record()
        nat.call("jfm_l2p", s, code, p, state.pos4.data_ptr(), N, leaf, topo.num_internal, state.locals.data_ptr(),
                 state.centers4.data_ptr(), state.G, state.perm.data_ptr(), acc.data_ptr(), None)
        if ev is not None:
            ev[3].record()
        nat.call("jfm_p2p", s, code, state.pos4.data_ptr(), N, topo.num_leaves, leaf, state.near_off.data_ptr(),
                 state.near_src.data_ptr(), eps * eps, state.G, state.perm.data_ptr(), acc.data_ptr(), None, flags, 1)
        if ev is not None:
            ev[4].record()

    for _ in range(args.warmup):
        run_step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        run_step(events[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1e3 * elapsed / args.steps
    stage = {"upward": [], "downward": [], "l2p": [], "p2p": []}
    for ev in events:
        stage["upward"].append(ev[0].elapsed_time(ev[1]))
        stage["downward"].append(ev[1].elapsed_time(ev[2]))
        stage["l2p"].append(ev[2].elapsed_time(ev[3]))
        stage["p2p"].append(ev[3].elapsed_time(ev[4]))
    stage_ms = {k: float(np.mean(v)) for k, v in stage.items()}
    log(f"{ms_per_step:.3f} ms/step; stages (ms) {json.dumps({k: round(v, 3) for k, v in stage_ms.items()})}")

    # ---- roofline of the dominant kernel (P2P, FP32 VALU) -------------------
    interactions = topo.p2p_interactions()
    p2p_flops = 20.0 * interactions
    p2p_s = stage_ms["p2p"] * 1e-3
    achieved = p2p_flops / p2p_s / 1e12
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"p2p_{args.workload}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "valu", "kernel": "p2p_kernel", "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS if precision == "fp32" else 78.6, "unit": "TFLOP/s",
                "frac": round(achieved / (PEAK_FP32_TFLOPS if precision == "fp32" else 78.6), 4),
                "traffic": traffic, "interactions_per_launch": int(interactions), "flops_per_interaction": 20,
                "avg_launch_ms": round(stage_ms["p2p"], 4)}

    # ---- force error against the FP64 direct sum ------------------------------
    err = None
    if args.error_targets > 0:
        rng = np.random.default_rng(12345)
        tgt = np.sort(rng.choice(N, size=min(args.error_targets, N), replace=False))
        a_full = torch.empty((N, 3), dtype=wdt, device=dev)
        fmm._evaluate_device(state, a_full, None)
        ref = direct_sum_gravitational_acceleration(x, m, softening=eps, target_indices=tgt)
        got = a_full[torch.as_tensor(tgt, device=dev)].to(torch.float64)
        rel = (torch.linalg.norm(got - ref, dim=1) / torch.linalg.norm(ref, dim=1)).cpu().numpy()
        err = {"median": float(np.median(rel)), "p99": float(np.percentile(rel, 99)), "max": float(rel.max()),
               "targets": int(tgt.size), "target_seed": 12345}
        log(f"force error vs FP64 direct sum: median {err['median']:.3e} p99 {err['p99']:.3e}")

    cpu = None
    if not args.no_cpu_baseline:
        cn = min(args.cpu_sample, N)
        rate, dt = cpu_baseline_sample(gen, p, leaf, theta, eps, cn)
        cpu = {"value": rate, "unit": "particles/s", "cores": 1, "kind": "port",
               "sample": f"float64 NumPy oracle FMM, full pipeline (tree, lists, P2M..P2P), {gen} N={cn}, "
                         f"p={p}, leaf={leaf}, theta={theta}, one process ({dt:.1f}s)",
               "cpu": platform_cpu()}
        log(f"cpu baseline {rate:.3e} particles/s on N={cn} ({dt:.1f}s)")

    value = N / (ms_per_step * 1e-3)
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
        "config": {"workload": f"{gen} N={N} p={p} leaf={leaf} theta={theta} softening={eps} {precision}, "
                               f"fixed-order plain FMM", "N": N, "seed": seed, "order": p, "leaf_size": leaf,
                   "theta": theta, "softening": eps, "parallelism": "single-gpu",
                   "mode": "plain fixed-order, compensated P2P"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "force_error": err,
        "stage_ms": stage_ms,
        "prepare_s": prep_s,
        "topology_build_s": topo.build_seconds,
        "far_pairs": topo.num_far_pairs,
        "m2l_gflop_per_step": m2l_flops(p) * topo.num_far_pairs / 1e9,
    }
    print(json.dumps(out), flush=True)
    return 0


def platform_cpu() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())
