"""Train the per-pair M2L order model on the GPU and save its weights.

Training data come from the FMM itself on the composite-galaxy workload
(Hernquist halo a=1 holding 80% of the equal-mass particles, exponential disc
R_d=0.3, z_d=0.03): features of every far pair from the order-6 expansions and
labels from order-8 translations, both computed by libjfm kernels; the MLP is
trained with libjfm's backward and Adam kernels.

    python -m jaccpot_amd.train_order [--n 2097152] [--seed 13] [--leaf 16] [--theta 0.7] [--steps 2000]
                                      [--gears 2,4,6] [--out PATH]

The default galaxy seed (13) differs from the seed of the learned-order
workload (3), so that workload is held out from the training data.
"""

from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

MODELS = Path(__file__).resolve().parent / "models"
DEFAULT_WEIGHTS = MODELS / "order_mlp_gears246.npz"


def weights_path(gears) -> Path:
    """Shipped weights for a gear triple, e.g. models/order_mlp_gears368.npz."""
    return MODELS / ("order_mlp_gears" + "".join(str(int(g)) for g in gears) + ".npz")


def train(n: int = 2_097_152, seed: int = 13, leaf: int = 16, theta: float = 0.7, steps: int = 2000,
          batch: int = 32768, lr: float = 3e-3, gears=(2, 4, 6), log=print):
    import torch

    from . import FastMultipoleMethod, ics
    from .learned import OrderModel, pair_features, pair_labels

    x, m = ics.galaxy(n, seed)
    fmm = FastMultipoleMethod(precision="fp32", theta=theta, softening=1e-3)
    t0 = time.perf_counter()
    st = fmm.prepare_state(x.astype(np.float32), m.astype(np.float32), leaf_size=leaf, max_order=max(gears))
    X = pair_features(st)
    T = pair_labels(st, gears)
    torch.cuda.synchronize()
    log(f"training data: {X.shape[0]} far pairs from galaxy N={n} leaf={leaf} theta={theta} "
        f"({time.perf_counter() - t0:.1f}s)")
    model = OrderModel(gears=tuple(int(g) for g in gears), seed=seed)
    t1 = time.perf_counter()
    hist = model.fit(X, T, steps=steps, batch=batch, lr=lr, log_every=max(1, steps // 10))
    torch.cuda.synchronize()
    for it, loss in hist:
        log(f"  step {it:5d} loss {loss:.4f}")
    resid = (model.predict(X) - T).abs().median().item()
    log(f"trained in {time.perf_counter() - t1:.1f}s; median |log10 residual| {resid:.3f}")
    return model, {"loss_history": hist, "median_abs_residual": resid, "pairs": int(X.shape[0]), "n": n,
                   "leaf": leaf, "theta": theta, "seed": seed, "gears": tuple(gears)}


def save(model, path: Path, meta: dict) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    d = model.state_dict()
    np.savez(path, master=d["master"], gears=d["gears"], feature_mean=d["feature_mean"],
             feature_std=d["feature_std"], meta_n=meta["n"], meta_leaf=meta["leaf"], meta_theta=meta["theta"],
             meta_residual=meta["median_abs_residual"], meta_pairs=meta["pairs"], meta_seed=meta["seed"])


def load(path: Path = DEFAULT_WEIGHTS):
    """An OrderModel with saved weights (float32 masters -> bfloat16 copies)."""
    from .learned import OrderModel

    with np.load(path, allow_pickle=False) as f:
        d = {k: f[k] for k in ("master", "gears", "feature_mean", "feature_std")}
    model = OrderModel(gears=tuple(int(g) for g in d["gears"]))
    model.load_state_dict(d)
    return model


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=2_097_152)
    ap.add_argument("--seed", type=int, default=13, help="galaxy seed of the training data")
    ap.add_argument("--leaf", type=int, default=16)
    ap.add_argument("--theta", type=float, default=0.7)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--gears", default="2,4,6", help="three increasing gears in 1..8, e.g. 3,6,8")
    ap.add_argument("--out", type=Path, default=None)
    a = ap.parse_args(argv)
    gears = tuple(int(g) for g in a.gears.split(","))
    model, meta = train(n=a.n, seed=a.seed, leaf=a.leaf, theta=a.theta, steps=a.steps, gears=gears,
                        log=lambda s: print(s, file=sys.stderr, flush=True))
    a.out = a.out or weights_path(gears)
    save(model, a.out, meta)
    print(a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
