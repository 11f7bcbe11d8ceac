"""GPU parity of the matrix-core M2L engine (float32 B_U rotations on
v_mfma_f32_4x4x1_16b_f32) against the float64 oracle and the VALU engine.

The engine changes only how the constant B_U blocks are applied, so:

* the float32 pipeline with ``m2l_engine="mfma"`` matches the float64 oracle
  to the float32 bar (1e-5 relative L2) at every order the per-target kernel
  runs (p = 3, 4, 6, 8);
* its M2L locals agree row by row with the VALU engine's to float32 rounding;
* the lane-per-pair (flat) kernel of the learned gears matches the float64
  C++ restatement with the same per-pair orders on the same topology.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fmm as ofmm
from oracle import ics

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _rel_rows(a, b):
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    return float((np.abs(a - b) / scale).max())


CASES = [
    ("uniform4096_leaf16_p4", "uniform", 4096, 16, 4, 0.6, 0.0),
    ("plummer8192_leaf32_p3", "plummer", 8192, 32, 3, 0.6, 1e-3),
    ("plummer8192_leaf64_p6", "plummer", 8192, 64, 6, 0.6, 1e-3),
    ("blobs4096_leaf16_p8", "blobs", 4096, 16, 8, 0.6, 1e-3),
    ("ragged77_leaf4_p8", "uniform", 77, 4, 8, 0.5, 0.0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_mfma_engine_pipeline_matches_oracle(case, require_gpu):
    from jaccpot_amd import FastMultipoleMethod

    name, gen, n, leaf, p, theta, eps = case
    print(f"[gpu] mfma M2L engine fp32 pipeline {name}", flush=True)
    x, m = getattr(ics, gen)(n, 1)
    x32, m32 = x.astype(np.float32), m.astype(np.float32)
    out = {}
    for engine in ("valu", "mfma"):
        fmm = FastMultipoleMethod(precision="fp32", theta=theta, softening=eps, m2l_engine=engine)
        st = fmm.prepare_state(x32, m32, leaf_size=leaf, max_order=p)
        acc = fmm.evaluate_prepared_state(st)
        torch.cuda.synchronize()
        out[engine] = (acc, st.locals.cpu().numpy()[:, : (p + 1) ** 2].astype(np.float64))
    ref = ofmm.fmm_accelerations(x32.astype(np.float64), m32.astype(np.float64), leaf_size=leaf, p=p, theta=theta,
                                 eps=eps)
    rel = np.linalg.norm(out["mfma"][0] - ref) / np.linalg.norm(ref)
    rel_valu = np.linalg.norm(out["valu"][0] - ref) / np.linalg.norm(ref)
    lv, lm = out["valu"][1], out["mfma"][1]
    nz = np.abs(lv).max(axis=1) > 0
    drow = _rel_rows(lm[nz], lv[nz]) if nz.any() else 0.0
    print(f"[gpu]   rel-L2 vs oracle: mfma {rel:.3e} valu {rel_valu:.3e}; locals mfma vs valu max row rel {drow:.3e}",
          flush=True)
    assert rel < 1e-5
    assert drow < 1e-4


def test_mfma_engine_flat_gears_match_cpu_restatement(require_gpu):
    """Learned gears (3, 6, 8) on Plummer N=65,536 (leaf 8, theta 0.85): the flat
    lane-per-pair kernels on the matrix-core engine against the float64 C++
    restatement with the same per-pair orders; the VALU engine beside it."""
    from jaccpot_amd import FastMultipoleMethod
    from oracle import cfmm

    if not cfmm.LIBRARY.exists():
        cfmm.build()
    print("[gpu] mfma M2L engine, learned gears (3,6,8) flat kernels vs C++ restatement", flush=True)
    x, m = ics.plummer(65_536, 7)
    x32, m32 = x.astype(np.float32), m.astype(np.float32)
    accs = {}
    for engine in ("valu", "mfma"):
        fmm = FastMultipoleMethod(precision="fp32", theta=0.85, softening=1e-3, adaptive_order=True,
                                  p_gears=(3, 6, 8), gear_tolerance=-1.5, m2l_engine=engine)
        st = fmm.prepare_state(x32, m32, leaf_size=8, max_order=8)
        accs[engine] = fmm.evaluate_prepared_state(st)
        torch.cuda.synchronize()
        orders = st.extra["gear_of_pair"]
        assert st.extra.get("gears") is not None and any(f is not None for f in st.extra["gears"].flat)
    counts = {int(g): int(np.sum(orders == g)) for g in (3, 6, 8)}
    ref = cfmm.evaluate(st.topology, x32.astype(np.float64), m32.astype(np.float64), p=8, eps=1e-3, orders=orders)
    rel = np.linalg.norm(accs["mfma"] - ref) / np.linalg.norm(ref)
    rel_valu = np.linalg.norm(accs["valu"] - ref) / np.linalg.norm(ref)
    print(f"[gpu]   gears {counts}; rel-L2 vs restatement: mfma {rel:.3e} valu {rel_valu:.3e}", flush=True)
    assert rel < 1e-5


def test_engine_switch_rejects_unknown_engine(require_gpu):
    from jaccpot_amd import _native as nat

    assert nat.call("jfm_m2l_engine") in nat.M2L_ENGINES.values()
    with pytest.raises(nat.JfmError):
        nat.call("jfm_set_m2l_engine", 7)
