"""Float64 direct summation on the GPU (libjfm ``jfm_direct_sum_f64``).

Counterpart of ``direct_sum_gravitational_acceleration`` (jaccpot/autodiff.py:12-98):
softened Newtonian accelerations with the self term masked.  Used to report
force errors of every FMM result against an FP64 direct sum.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat


def direct_sum_gravitational_acceleration(positions, masses, *, G: float = 1.0, softening: float = 0.0,
                                          target_indices=None, return_potential: bool = False):
    """FP64 direct sum at ``target_indices`` (default: all particles).

    Accepts numpy arrays or torch tensors; returns float64 torch tensors on the
    GPU (``acc [T,3]`` and optionally ``pot [T]``).
    """
    if not torch.cuda.is_available():
        raise RuntimeError("direct_sum_gravitational_acceleration needs a ROCm GPU")
    dev = torch.device("cuda", torch.cuda.current_device())
    x = torch.as_tensor(positions).to(dev, torch.float64).contiguous()
    m = torch.as_tensor(masses).to(dev, torch.float64).contiguous()
    n = x.shape[0]
    if x.ndim != 2 or x.shape[1] != 3 or m.shape != (n,):
        raise ValueError("positions must be (N,3) and masses (N,)")
    if target_indices is None:
        tgt = torch.arange(n, dtype=torch.int32, device=dev)
    else:
        tgt = torch.as_tensor(np.asarray(target_indices) if not isinstance(target_indices, torch.Tensor)
                              else target_indices).to(dev, torch.int32).contiguous()
        if tgt.numel() and (int(tgt.min()) < 0 or int(tgt.max()) >= n):
            raise ValueError("target_indices out of range")
    nt = int(tgt.numel())
    pos4 = torch.empty((n, 4), dtype=torch.float64, device=dev)
    ident = torch.arange(n, dtype=torch.int32, device=dev)
    s = nat.current_stream()
    nat.call("jfm_gather_sorted", s, nat.JFM_F64, nat.JFM_F64, x.data_ptr(), m.data_ptr(), ident.data_ptr(), n,
             pos4.data_ptr())
    acc = torch.empty((nt, 3), dtype=torch.float64, device=dev)
    pot = torch.empty((nt,), dtype=torch.float64, device=dev) if return_potential else None
    if nt:
        scratch = torch.empty(int(nat.load().jfm_direct_scratch_doubles(nt)), dtype=torch.float64, device=dev)
        nat.call("jfm_direct_sum_f64", s, pos4.data_ptr(), n, tgt.data_ptr(), nt, float(softening) ** 2, float(G),
                 scratch.data_ptr(), acc.data_ptr(), pot.data_ptr() if pot is not None else None)
    if return_potential:
        return acc, pot
    return acc
