"""The allocation-free direct sum (jfm_direct_sum_f64_static) against the
caller-scratch entry point (jfm_direct_sum_f64), both called through the C ABI.

Both run the same two kernels in the same order, so the results are equal bit
for bit.  Two calls in a row, the second with fewer targets, show that the
per-GPU scratch array left by one call does not leak into the next; a call
over the maximum target count is refused on the host.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _lib():
    from jaccpot_amd import _native as nat

    lib = nat.load()
    fn = lib.jfm_direct_sum_f64_static
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                   ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    mx = lib.jfm_direct_sum_f64_static_max_targets
    mx.restype = ctypes.c_int32
    mx.argtypes = []
    return nat, lib


def _pos4(n, seed):
    from oracle import ics

    x, m = ics.plummer(n, seed)
    return torch.as_tensor(np.concatenate([x, m[:, None]], axis=1), dtype=torch.float64, device="cuda").contiguous()


def _with_caller_scratch(nat, lib, s, pos4, tgt, eps2, G):
    nt = int(tgt.numel())
    acc = torch.empty((nt, 3), dtype=torch.float64, device="cuda")
    pot = torch.empty((nt,), dtype=torch.float64, device="cuda")
    scratch = torch.empty(int(lib.jfm_direct_scratch_doubles(nt)), dtype=torch.float64, device="cuda")
    nat.call("jfm_direct_sum_f64", s, pos4.data_ptr(), pos4.shape[0], tgt.data_ptr(), nt, eps2, G,
             scratch.data_ptr(), acc.data_ptr(), pot.data_ptr())
    return acc, pot


def _static(nat, lib, s, pos4, tgt, eps2, G):
    nt = int(tgt.numel())
    acc = torch.empty((nt, 3), dtype=torch.float64, device="cuda")
    pot = torch.empty((nt,), dtype=torch.float64, device="cuda")
    rc = lib.jfm_direct_sum_f64_static(s, pos4.data_ptr(), pos4.shape[0], tgt.data_ptr(), nt, eps2, G,
                                       acc.data_ptr(), pot.data_ptr())
    nat.check(rc, "jfm_direct_sum_f64_static")
    return acc, pot


def test_static_scratch_direct_sum_matches_caller_scratch(require_gpu):
    nat, lib = _lib()
    s = nat.current_stream()
    maxt = int(lib.jfm_direct_sum_f64_static_max_targets())
    assert maxt == 4096
    # first call: the maximum target count; second call: fewer targets, other particles
    for n, nt, seed, eps2, G in ((20_000, maxt, 1, 1e-6, 1.0), (7_000, 1_000, 2, 0.0, 4.3e-3)):
        print(f"[gpu] static-scratch direct sum n={n} targets={nt}", flush=True)
        pos4 = _pos4(n, seed)
        rng = np.random.default_rng(seed)
        tgt = torch.as_tensor(np.sort(rng.choice(n, nt, replace=False)).astype(np.int32), device="cuda")
        a0, p0 = _with_caller_scratch(nat, lib, s, pos4, tgt, eps2, G)
        a1, p1 = _static(nat, lib, s, pos4, tgt, eps2, G)
        torch.cuda.synchronize()
        assert torch.equal(a0, a1) and torch.equal(p0, p1)
        assert torch.isfinite(a1).all() and float(a1.abs().max()) > 0.0


def test_static_scratch_direct_sum_refuses_too_many_targets(require_gpu):
    nat, lib = _lib()
    s = nat.current_stream()
    maxt = int(lib.jfm_direct_sum_f64_static_max_targets())
    pos4 = _pos4(maxt + 1, 3)
    tgt = torch.arange(maxt + 1, dtype=torch.int32, device="cuda")
    out = torch.empty((maxt + 1, 3), dtype=torch.float64, device="cuda")
    rc = lib.jfm_direct_sum_f64_static(s, pos4.data_ptr(), maxt + 1, tgt.data_ptr(), maxt + 1, 0.0, 1.0,
                                       out.data_ptr(), None)
    assert rc == nat.JFM_ERR_ARG if hasattr(nat, "JFM_ERR_ARG") else rc == 1
    assert b"exceed the maximum" in lib.jfm_last_error()
