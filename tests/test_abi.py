"""C ABI surface of libjfm.so (CPU): loads, exports every declared symbol,
rejects bad arguments before touching the GPU, and its host-compiled operator
math agrees with the oracle."""

from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import harmonics as H

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "jfm.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jfm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from jaccpot_amd import _native as nat

    lib = nat.load()
    declared = _declared()
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", str(nat.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(jfm_[a-z0-9_]+)", out))
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        assert getattr(lib, s) is not None
    assert lib.jfm_abi_version() == nat.ABI_VERSION


def test_binding_covers_every_declared_symbol():
    from jaccpot_amd import _native as nat

    assert sorted(nat._SIGNATURES) == _declared()


def test_coeff_stride_and_supported_orders():
    from jaccpot_amd import _native as nat

    lib = nat.load()
    assert lib.jfm_coeff_stride(4, nat.JFM_F32) == 28 and lib.jfm_coeff_stride(6, nat.JFM_F32) == 52
    assert lib.jfm_coeff_stride(4, nat.JFM_F64) == 26 and lib.jfm_coeff_stride(6, nat.JFM_F64) == 50
    assert lib.jfm_order_supported(8, nat.JFM_F32) == 1 and lib.jfm_order_supported(9, nat.JFM_F32) == 0
    assert lib.jfm_coeff_stride(4, 7) == -1


def test_device_entry_points_reject_bad_arguments_without_launching():
    from jaccpot_amd import _native as nat

    lib = nat.load()
    assert lib.jfm_p2m(None, nat.JFM_F32, 4, None, 10, 3, 4, 2, None, None) == 1  # NULL buffers
    assert lib.jfm_p2m(None, nat.JFM_F32, 4, 8, 10, 2, 4, 1, 8, 8) == 1  # L != ceil(N/leaf)
    assert b"inconsistent" in lib.jfm_last_error()
    assert lib.jfm_m2l(None, nat.JFM_F32, 4, None, -1, None, None, None, None, None) == 1
    assert lib.jfm_m2l(None, nat.JFM_F32, 4, None, 0, None, None, None, None, None) == 0  # empty is a no-op
    assert lib.jfm_p2p(None, nat.JFM_F32, 8, 10, 3, 4, 8, 8, -1.0, 1.0, None, 8, None, 0, 0) == 1  # eps2 < 0
    assert lib.jfm_p2p(None, nat.JFM_F32, 8, 1000, 2, 500, 8, 8, 0.0, 1.0, None, 8, None, 0, 0) == 1  # leaf > 256
    assert lib.jfm_evaluate(None, nat.JFM_F32, 4, 8, 10, 3, 4, 2, 8, 8, 8, 8, 0.0, 1.0, None, 8, None, 0) == 1
    assert lib.jfm_direct_sum_f64(None, None, 0, None, 1, 0.0, 1.0, None, None, None) == 1
    assert lib.jfm_direct_scratch_doubles(4096) == 256 * 4 * 4096
    assert lib.jfm_host_operator(9, 4, 1, ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8)) == 1


def test_tables_export_matches_oracle_bu():
    from jaccpot_amd import _native as nat

    lib = nat.load()
    p = 8
    size = sum((2 * l + 1) ** 2 for l in range(p + 1))
    buf = np.zeros(size)
    assert lib.jfm_tables_bu(p, buf.ctypes.data, size) == 0
    off = 0
    for ell in range(p + 1):
        w = 2 * ell + 1
        np.testing.assert_allclose(buf[off : off + w * w].reshape(w, w), H.B_U(ell), atol=1e-14)
        off += w * w
    assert lib.jfm_tables_bu(p, buf.ctypes.data, size - 1) == 1


def _host_op(op, p, vin, delta, dtype):
    from jaccpot_amd import _native as nat

    out = np.zeros(max(H.sh_size(p), 4))
    vin = np.ascontiguousarray(vin, np.float64)
    d = np.ascontiguousarray(delta, np.float64)
    nat.check(nat.load().jfm_host_operator(op, p, dtype, vin.ctypes.data, d.ctypes.data, out.ctypes.data), "op")
    return out[: H.sh_size(p)] if op != 4 else out[:4]


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6, 7, 8])
def test_kernel_operator_math_matches_oracle(p):
    """The device templates instantiated on the host (float64 and float32)."""
    rng = np.random.default_rng(100 + p)
    C = H.sh_size(p)
    for trial in range(3):
        v = rng.normal(size=C)
        d = rng.normal(size=3) * (1.0 + trial)
        for dtype, tol in ((1, 1e-12), (0, 2e-5)):
            for op, fn in ((0, H.m2m), (1, H.m2l), (2, H.l2l)):
                ref = fn(v, d, p)
                got = _host_op(op, p, v, d, dtype)
                assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (op, dtype)
            ref = H.p2m(d[None], np.array([0.7]), p)
            assert np.abs(_host_op(3, p, [0.7], d, dtype) - ref).max() <= tol * np.abs(ref).max()
            phi, grad = H.l2p(v, d, p)
            got = _host_op(4, p, v, d, dtype)
            assert abs(got[0] - phi) <= tol * max(1.0, abs(phi)) * 10
            assert np.abs(got[1:] - grad).max() <= tol * max(1.0, np.abs(grad).max()) * 10


def test_kernel_operator_degenerate_directions():
    """z-aligned and zero displacements (rho = 0 and r = 0 branches)."""
    p = 6
    v = np.random.default_rng(3).normal(size=H.sh_size(p))
    for d in ([0.0, 0.0, 2.5], [0.0, 0.0, -1.5], [0.0, 0.0, 0.0], [1e-9, 0.0, 3.0]):
        d = np.array(d)
        for op, fn in ((0, H.m2m), (2, H.l2l)):
            np.testing.assert_allclose(_host_op(op, p, v, d, 1), fn(v, d, p), atol=1e-12 * np.abs(v).max() * 10)
        if np.linalg.norm(d) > 0:
            ref = H.m2l(v, d, p)
            np.testing.assert_allclose(_host_op(1, p, v, d, 1), ref, atol=1e-12 * np.abs(ref).max())
