"""ctypes wrapper of ``cfmm.cpp``, the C++/OpenMP restatement of the FMM numerics.

Oracle / test infrastructure only (see ``oracle/__init__.py``): it checks the
HIP path at sizes the NumPy pipeline cannot reach, and it is the CPU baseline
of the benchmark (the reference's plain fixed-order pipeline timed on host
cores).  ``build()`` compiles ``oracle/lib/libcfmm.so`` with g++ (no HIP, no
torch).

Pinned by: tests/test_cfmm.py, against ``oracle.fmm.fmm_accelerations`` in
float64 on the same topology (rel-L2 < 1e-13 on accelerations and
potentials, p = 2..8), and the float32 build against the float64 one.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from . import harmonics as H

HERE = Path(__file__).resolve().parent
SOURCE = HERE / "cfmm.cpp"
LIBRARY = HERE / "lib" / "libcfmm.so"
CXX = os.environ.get("CXX", "g++")
# x86-64-v3 (AVX2 + FMA): built here, run on whichever x86 host the library travels to.
FLAGS = ["-O3", "-march=x86-64-v3", "-fno-math-errno", "-fopenmp", "-std=c++17", "-shared", "-fPIC"]

_lib = None


def build(verbose: bool = False) -> Path:
    """Compile libcfmm.so (always; a few seconds)."""
    LIBRARY.parent.mkdir(exist_ok=True)
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [CXX, *FLAGS, "-o", str(tmp), str(SOURCE)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIBRARY)
    return LIBRARY


def load():
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise ImportError(f"{LIBRARY} is not built: python -c 'import oracle.cfmm as c; c.build()'")
        lib = ctypes.CDLL(str(LIBRARY))
        i32, f64, vp = ctypes.c_int32, ctypes.c_double, ctypes.c_void_p
        lib.cfmm_version.restype = i32
        lib.cfmm_max_threads.restype = i32
        lib.cfmm_evaluate.restype = i32
        lib.cfmm_evaluate.argtypes = [i32, i32, i32, vp, vp, vp, i32, i32] + [vp] * 12 + [f64, f64, i32, vp, vp, vp]
        _lib = lib
    return _lib


def bu_blocks(p: int) -> np.ndarray:
    """Dense B_U blocks for degrees 0..p, row-major, concatenated."""
    return np.ascontiguousarray(np.concatenate([H.B_U(l).ravel() for l in range(p + 1)]), dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.int32)


def evaluate(topo, positions, masses, *, p: int, eps: float = 0.0, G: float = 1.0, precision: str = "fp64",
             orders=None, threads: int = 0, return_potential: bool = False, timings: dict | None = None):
    """Fixed-order FMM accelerations (and potentials) in input order on ``topo``.

    ``topo`` is an ``oracle.topology.Topology`` or any object with the same
    fields (perm, node_range, left, right, parent, depth, center, far/near CSR).
    ``precision`` selects the arithmetic type ("fp64" or "fp32"); ``orders``
    (optional, [K] in far-list order) gives each far pair's M2L order <= p.
    """
    lib = load()
    x = np.ascontiguousarray(np.asarray(positions, np.float64))
    m = np.ascontiguousarray(np.asarray(masses, np.float64))
    n = x.shape[0]
    arrs = dict(
        perm=_i32(topo.perm), node_range=_i32(topo.node_range), left=_i32(topo.left), right=_i32(topo.right),
        parent=_i32(topo.parent), depth=_i32(topo.depth), far_off=_i32(topo.far_off), far_src=_i32(topo.far_src),
        near_off=_i32(topo.near_off), near_src=_i32(topo.near_src),
    )
    center = np.ascontiguousarray(np.asarray(topo.center, np.float64))
    bu = bu_blocks(p)
    od = None if orders is None else np.ascontiguousarray(np.asarray(orders), dtype=np.int8)
    if od is not None and od.shape != (np.asarray(topo.far_src).shape[0],):
        raise ValueError("orders must have one entry per far pair")
    acc = np.zeros((n, 3))
    pot = np.zeros(n) if return_potential else None
    sec = np.zeros(6)

    def ptr(a):
        return a.ctypes.data if a is not None and a.size else None

    rc = lib.cfmm_evaluate(
        {"fp64": 0, "fp32": 1}[precision], int(p), n, ptr(x), ptr(m), ptr(arrs["perm"]), int(topo.num_internal),
        int(topo.num_leaves), ptr(arrs["node_range"]), ptr(arrs["left"]), ptr(arrs["right"]), ptr(arrs["parent"]),
        ptr(arrs["depth"]), ptr(center), ptr(arrs["far_off"]), ptr(arrs["far_src"]), ptr(arrs["near_off"]),
        ptr(arrs["near_src"]), ptr(od), ptr(bu), float(G), float(eps), int(threads), ptr(acc), ptr(pot), ptr(sec),
    )
    if rc != 0:
        raise ValueError(f"cfmm_evaluate failed with status {rc}")
    if timings is not None:
        timings.update(zip(("gather", "p2m", "m2m", "m2l", "l2l", "l2p_p2p"), sec.tolist()))
    return (acc, pot) if return_potential else acc
