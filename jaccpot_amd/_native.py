"""ctypes binding of libjfm.so (the C ABI declared in include/jfm.h).

torch is imported first on purpose: libjfm.so resolves ``libamdhip64.so.7`` by
soname, so loading it after torch binds it to the HIP runtime torch already
mapped (one runtime per process, shared streams and allocations).

The library is required: there is no CPU fallback anywhere in this package.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_void_p
from pathlib import Path

import torch  # noqa: F401  (must precede loading libjfm.so, see module docstring)

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libjfm.so"

JFM_OK = 0
JFM_F32 = 0
JFM_F64 = 1
EVAL_NEAR = 1
EVAL_FAR = 2
EVAL_POTENTIAL = 4
EVAL_COMPENSATED = 8
MAX_ORDER = 8
M2L_VALU = 0
M2L_MFMA = 1
M2L_ENGINES = {"valu": M2L_VALU, "mfma": M2L_MFMA}
ABI_VERSION = 13

_i32p = POINTER(c_int32)
_f64p = POINTER(c_double)

# name -> (restype, argtypes)
_SIGNATURES = {
    "jfm_last_error": (c_char_p, []),
    "jfm_abi_version": (c_int32, []),
    "jfm_coeff_stride": (c_int32, [c_int32, c_int32]),
    "jfm_order_supported": (c_int32, [c_int32, c_int32]),
    "jfm_topology_build": (c_int32, [c_void_p, c_void_p, c_int64, c_int32, c_double, c_void_p, c_int32,
                                     POINTER(c_void_p)]),
    "jfm_topology_info": (c_int32, [c_void_p, c_void_p]),
    "jfm_topology_export": (c_int32, [c_void_p] + [c_void_p] * 13),
    "jfm_topology_schedules": (c_int32, [c_void_p] + [c_void_p] * 5),
    "jfm_topology_free": (None, [c_void_p]),
    "jfm_tree_bounds_scratch_doubles": (c_int32, []),
    "jfm_tree_bounds": (c_int32, [c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p]),
    "jfm_tree_morton": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "jfm_sort_scratch_bytes": (c_int32, [c_int64, c_int32, POINTER(c_int64)]),
    "jfm_sort_u64": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_int64]),
    "jfm_tree_radix": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_tree_levels": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p]),
    "jfm_tree_leaf_geometry": (c_int32, [c_void_p, c_int32, c_int32] + [c_void_p] * 6),
    "jfm_tree_level_geometry": (c_int32, [c_void_p, c_int32] + [c_void_p] * 9),
    "jfm_tree_radius": (c_int32, [c_void_p, c_int32, c_int64] + [c_void_p] * 6),
    "jfm_tree_walk_classify": (c_int32, [c_void_p, c_int32, c_void_p, c_int64, c_void_p, c_void_p, c_double,
                                         c_void_p, c_void_p]),
    "jfm_tree_walk_emit": (c_int32, [c_void_p, c_int32, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_int64, c_int64, c_void_p, c_void_p, c_void_p]),
    "jfm_scan_scratch_bytes": (c_int32, [c_int64, POINTER(c_int64)]),
    "jfm_scan_i32": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64]),
    "jfm_keys_to_csr": (c_int32, [c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p]),
    "jfm_key_bits": (c_int32, [ctypes.c_uint64]),
    "jfm_pair_rows": (c_int32, [c_void_p, c_int32, c_void_p, c_int64, c_void_p]),
    "jfm_gear_keys": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_gather_pairs": (c_int32, [c_void_p, c_int64] + [c_void_p] * 5),
    "jfm_pair_keys": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "jfm_nonempty_rows": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    "jfm_flat_spans": (c_int32, [c_void_p, c_int64] + [c_void_p] * 5),
    "jfm_tables_bu": (c_int32, [c_int32, c_void_p, c_int64]),
    "jfm_gather_sorted": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "jfm_p2m": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p,
                          c_void_p]),
    "jfm_m2m_level": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p]),
    "jfm_m2l": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                          c_void_p]),
    "jfm_m2l_gear": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p]),
    "jfm_m2l_flat_scratch_size": (c_int64, [c_int32, c_int32]),
    "jfm_m2l_flat": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    "jfm_set_m2l_engine": (c_int32, [c_int32]),
    "jfm_m2l_engine": (c_int32, []),
    "jfm_l2l_level": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    "jfm_m2m_top": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_l2l_top": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_l2p": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p,
                          c_double, c_void_p, c_void_p, c_void_p, c_int32]),
    "jfm_l2p_combine": (c_int32, [c_void_p, c_int32, c_int32, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_p2p": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_double,
                          c_double, c_void_p, c_void_p, c_void_p, c_int32, c_int32]),
    "jfm_p2p_items": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_double, c_double,
                                c_void_p, c_void_p, c_void_p, c_int32, c_int32]),
    "jfm_p2p_reduce": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_double, c_double, c_void_p, c_void_p, c_void_p, c_int32,
                                 c_int32]),
    "jfm_evaluate": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_double, c_double, c_void_p, c_void_p, c_void_p,
                               c_int32]),
    "jfm_direct_scratch_doubles": (c_int64, [c_int32]),
    "jfm_p2p_adjoint_items": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32,
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                                        c_double, c_double, c_void_p]),
    "jfm_p2p_adjoint_reduce": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_double, c_double, c_void_p]),
    "jfm_p2p_adjoint": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_void_p,
                                  c_void_p, c_double, c_double, c_void_p]),
    "jfm_leaf_adjoint_a": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                     c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "jfm_center_adjoint": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_mutual_rungs": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_mutual_rungs_level": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    "jfm_mutual_p2p": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_int32, c_double, c_double, c_void_p, c_void_p]),
    "jfm_mutual_p2p_items": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_double, c_double,
                                       c_void_p, c_void_p]),
    "jfm_m2l_flat_weighted": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                                        c_void_p, c_void_p]),
    "jfm_mutual_combine": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "jfm_p2p_adjoint_weighted": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                           c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_double, c_double,
                                           c_void_p]),
    "jfm_p2p_adjoint_weighted_items": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                                 c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                                                 c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_double, c_double,
                                                 c_void_p]),
    "jfm_center_adjoint_weighted": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p]),
    "jfm_gear_node_max": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p]),
    "jfm_level_path_max": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p]),
    "jfm_level_subtree_min": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    "jfm_select_gears_budget": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_float,
                                          c_void_p, c_void_p, c_void_p]),
    "jfm_center_adjoint_gears": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    "jfm_com_chain": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p]),
    "jfm_leaf_adjoint_b": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_int32, c_int32, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_pair_features": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_double, c_void_p]),
    "jfm_pair_labels": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p]),
    "jfm_mlp_num_params": (c_int32, []),
    "jfm_mlp_backward_scratch_floats": (c_int64, [c_int32, c_int32]),
    "jfm_mlp_forward": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_mlp_forward_mfma": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_mlp_backward": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_mlp_adam": (c_int32, [c_void_p, c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jfm_mlp_to_bf16": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "jfm_select_gears": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_float,
                                   c_void_p, c_void_p]),
    "jfm_host_operator": (c_int32, [c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p]),
    "jfm_direct_sum_f64_static_max_targets": (c_int32, []),
    "jfm_direct_sum_f64_static": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_double, c_double,
                                            c_void_p, c_void_p]),
    "jfm_direct_sum_f64": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_double, c_double, c_void_p,
                                     c_void_p, c_void_p]),
}

_lib = None


class JfmError(RuntimeError):
    """A libjfm entry point returned an error code."""


def load() -> ctypes.CDLL:
    """Load libjfm.so once (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m jaccpot_amd.build` "
            "(there is no fallback implementation)"
        )
    lib = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_LOCAL)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.jfm_abi_version() != ABI_VERSION:
        raise ImportError(f"libjfm.so ABI {lib.jfm_abi_version()} != expected {ABI_VERSION}; rebuild it")
    _lib = lib
    return lib


def check(rc: int, name: str) -> None:
    if rc != JFM_OK:
        msg = load().jfm_last_error()
        raise JfmError(f"{name} failed ({rc}): {msg.decode() if msg else ''}")


def call(name: str, *args) -> int:
    """Call an entry point and raise JfmError on a non-zero status."""
    rc = getattr(load(), name)(*args)
    if name not in ("jfm_direct_scratch_doubles", "jfm_coeff_stride", "jfm_order_supported", "jfm_abi_version",
                    "jfm_mlp_num_params", "jfm_mlp_backward_scratch_floats", "jfm_tree_bounds_scratch_doubles",
                    "jfm_key_bits", "jfm_m2l_engine", "jfm_direct_sum_f64_static_max_targets"):
        check(rc, name)
    return rc


def ptr(t) -> int | None:
    """Raw data pointer of a torch tensor / numpy array (None passes NULL)."""
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return t.data_ptr()
    return t.ctypes.data


def dtype_code(dtype) -> int:
    if dtype in (torch.float32, "float32", "fp32"):
        return JFM_F32
    if dtype in (torch.float64, "float64", "fp64"):
        return JFM_F64
    import numpy as np

    if np.dtype(dtype) == np.float32:
        return JFM_F32
    if np.dtype(dtype) == np.float64:
        return JFM_F64
    raise ValueError(f"unsupported dtype {dtype}")


def coeff_stride(p: int, dtype) -> int:
    return int(load().jfm_coeff_stride(p, dtype_code(dtype)))


def current_stream() -> int | None:
    """hipStream_t of torch's current stream on the current device."""
    return torch.cuda.current_stream().cuda_stream
