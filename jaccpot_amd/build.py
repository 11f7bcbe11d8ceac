"""In-tree build of libjfm.so (HIP kernels for gfx950 + host topology code).

Steps: compile and run the B_U table generator (float64, host), compile every
``csrc/*.hip`` for gfx950 with hipcc and every ``csrc/*.cpp`` with g++, link
``jaccpot_amd/_lib/libjfm.so``.  Invoked by ``__graft_entry__.build()`` and by
``python -m jaccpot_amd.build``.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = ROOT / "_build"
GEN = BUILD / "gen"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libjfm.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 instructions are disabled for every device compile; hipcc -O3
# would otherwise form v_pk_* from ordinary float code.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-Wno-unknown-warning-option",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall"]
# Per-source device flags.  m2l.hip: matrix-instruction results in ordinary
# VGPRs (the B_U products feed VALU code directly; in AGPRs every result
# would need a v_accvgpr_read first).
HIP_FILE_FLAGS = {"m2l.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(map(str, cmd))}\n{res.stdout}\n{res.stderr}")
    return res


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile everything and link libjfm.so; returns the library path."""
    BUILD.mkdir(exist_ok=True)
    GEN.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    obj = BUILD / "obj"
    obj.mkdir(exist_ok=True)

    gen_exe = BUILD / "gen_tables"
    _run([CXX, *CXX_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", CSRC / "gen_tables.cpp", "-o", gen_exe], verbose)
    _run([gen_exe, GEN / "tables_gen.h"], verbose)

    hip_srcs = sorted(CSRC.glob("*.hip"))
    cpp_srcs = sorted(p for p in CSRC.glob("*.cpp") if p.name != "gen_tables.cpp")
    inc = [f"-I{INCLUDE}", f"-I{CSRC}", f"-I{GEN}"]
    cmds = []
    objs = []
    for src in hip_srcs:
        o = obj / (src.stem + ".hip.o")
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, *HIP_FILE_FLAGS.get(src.name, []), *inc, "-c", src, "-o", o])
    for src in cpp_srcs:
        o = obj / (src.stem + ".cpp.o")
        objs.append(o)
        cmds.append([CXX, *CXX_FLAGS, *inc, "-c", src, "-o", o])
    jobs = jobs or min(len(cmds), os.cpu_count() or 4, 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_run, c, verbose) for c in cmds]:
            fut.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-pthread"], verbose)
    shutil.move(str(tmp), str(LIB))
    return LIB


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
