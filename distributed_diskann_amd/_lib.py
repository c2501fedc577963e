"""ctypes binding of libddg.so (include/ddg.h).  Loading fails loudly: there
is no CPU fallback for any entry point."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libddg.so")

DDG_OK, DDG_EINVAL, DDG_ERANGE, DDG_EDEGREE, DDG_EOOM, DDG_EDEVICE, DDG_EIO = range(7)
STATUS_NAMES = {0: "DDG_OK", 1: "DDG_EINVAL", 2: "DDG_ERANGE", 3: "DDG_EDEGREE", 4: "DDG_EOOM",
                5: "DDG_EDEVICE", 6: "DDG_EIO"}
DTYPES = {"uint8": 0, "int8": 1, "float32": 2}
METRICS = {"l2": 0, "mips": 1}
TRAV = {"native": 0, "bf16": 1, "fp8": 2}

# every symbol include/ddg.h declares
EXPORTS = [
    "ddg_create", "ddg_destroy", "ddg_last_error", "ddg_graph_set", "ddg_graph_get", "ddg_build",
    "ddg_batch_insert", "ddg_robust_prune", "ddg_search", "ddg_search_device", "ddg_insert_search",
    "ddg_ground_truth", "ddg_sort_rows", "ddg_size", "ddg_max_degree", "ddg_dims", "ddg_aligned_dims",
    "ddg_last_profile", "ddg_prepare_traversal", "ddg_distances", "ddg_autotune_search", "ddg_beam_search",
    "ddg_graph_set_rows", "ddg_graph_get_rows", "ddg_search_device_noalloc", "ddg_quantized_mips_distances",
    "ddg_search_shape",
]


class DdgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class BuildParamsC(C.Structure):
    _fields_ = [("R", C.c_uint32), ("L", C.c_uint32), ("alpha", C.c_double), ("two_pass", C.c_int),
                ("start_point", C.c_uint32), ("base", C.c_double), ("max_fraction", C.c_double),
                ("max_build", C.c_uint64), ("final_sort", C.c_int), ("bsp", C.c_int)]


class QueryParamsC(C.Structure):
    _fields_ = [("k", C.c_uint32), ("beam", C.c_uint32), ("cut", C.c_double), ("limit", C.c_uint64),
                ("degree_limit", C.c_uint32), ("trav", C.c_int), ("rerank", C.c_uint32)]


_lib = None


def lib():
    """Loads libddg.so (built by distributed_diskann_amd.build); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m distributed_diskann_amd.build` "
                          "(there is no CPU fallback)")
    # torch's wheel carries its own libamdhip64 / libhsa-runtime64 under the
    # same sonames as /opt/rocm's.  Loaded after libddg.so, it would be a
    # second HIP runtime in the process (torch then may find no device);
    # loaded first, libddg.so binds to it and the process has one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P, vp = C.POINTER, C.c_void_p
    u32, u64, i32, dbl = C.c_uint32, C.c_uint64, C.c_int, C.c_double
    sig = {
        "ddg_create": ([i32, i32, u64, u32, u32, vp, u32, P(vp)], i32),
        "ddg_destroy": ([vp], None),
        "ddg_last_error": ([vp], C.c_char_p),
        "ddg_graph_set": ([vp, vp], i32),
        "ddg_graph_get": ([vp, vp], i32),
        "ddg_build": ([vp, P(BuildParamsC), vp, P(dbl)], i32),
        "ddg_batch_insert": ([vp, vp, u64, dbl, u32, u32, i32, dbl, dbl], i32),
        "ddg_robust_prune": ([vp, u64, vp, vp, vp, vp, dbl, i32, u32, vp, vp], i32),
        "ddg_search": ([vp, vp, u64, vp, vp, u32, P(QueryParamsC), vp, vp, vp, vp, vp], i32),
        "ddg_search_device": ([vp, vp, u64, vp, vp, u32, P(QueryParamsC), vp, u32, vp, vp, vp, vp], i32),
        "ddg_insert_search": ([vp, vp, u64, u32, u32, u32, vp, vp, vp, vp], i32),
        "ddg_ground_truth": ([vp, vp, u64, u32, vp, vp], i32),
        "ddg_sort_rows": ([vp], i32),
        "ddg_size": ([vp], u64),
        "ddg_max_degree": ([vp], u32),
        "ddg_dims": ([vp], u32),
        "ddg_aligned_dims": ([vp], u32),
        "ddg_last_profile": ([vp, vp], i32),
        "ddg_prepare_traversal": ([vp, i32, P(dbl)], i32),
        "ddg_distances": ([vp, i32, vp, u64, vp, u32, vp], i32),
        "ddg_autotune_search": ([vp, vp, u64, P(QueryParamsC), vp, vp], i32),
        "ddg_beam_search": ([vp, vp, u64, vp, vp, u32, P(QueryParamsC), u32, vp, vp, vp, u32, vp, vp, vp, vp], i32),
        "ddg_graph_set_rows": ([vp, vp, u64, vp], i32),
        "ddg_quantized_mips_distances": ([i32, u32, C.c_float, C.c_float, vp, i32, u64, u32, vp, u64, u32, vp], i32),
        "ddg_search_device_noalloc": ([vp, vp, u64, vp, vp, u32, P(QueryParamsC), vp, u32, vp, vp, vp, vp], i32),
        "ddg_graph_get_rows": ([vp, vp, u64, vp], i32),
        "ddg_search_shape": ([vp, P(QueryParamsC), vp], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(status, handle=None):
    if status != DDG_OK:
        msg = lib().ddg_last_error(handle)
        raise DdgError(status, msg.decode() if msg else "")
