"""ctypes binding of libeac_hip.so plus the small torch plumbing the shims share.

The library is the product: every call below goes to a HIP kernel on the GPU.
If the shared library is missing or no GPU is visible, calls raise
`EACUnavailable` instead of silently computing something on the CPU.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).with_name("libeac_hip.so")

EAC_OK = 0
EAC_ERR_ARG = -1
EAC_ERR_HIP = -2
EAC_ERR_WORKSPACE = -3
EAC_ERR_KEY_WIDTH = -4
EAC_ERR_NOT_FOUND = -5
EAC_ERR_UNSUPPORTED = -6
EAC_ERR_VALUE = -7

W_I32, W_F32, W_U16, W_U8, W_I64, W_I16, W_U32 = 0, 1, 2, 3, 4, 5, 6

_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_sz = C.c_size_t

# name -> (restype, argtypes). Every symbol declared in include/eac_hip.h.
SIGNATURES = {
    "eac_last_error": (C.c_char_p, []),
    "eac_abi_version": (C.c_int, []),
    "eac_profile_enable": (C.c_int, [_i32]),
    "eac_profile_read": (C.c_int, [_i32, _p, _p, _p]),
    "eac_scan_workspace_bytes": (_sz, [_i64, _i32]),
    "eac_scan_exclusive_i32": (C.c_int, [_p, _p, _i64, _i32, _p, _p, _sz, _p]),
    "eac_scan_exclusive_i64": (C.c_int, [_p, _p, _i64, _i64, _p, _p, _sz, _p]),
    "eac_scan_exclusive_i32_i64": (C.c_int, [_p, _p, _i64, _i64, _p, _p, _sz, _p]),
    "eac_boruvka_workspace_bytes": (_sz, [_i32, _i64]),
    "eac_boruvka_mst": (C.c_int, [_p, _p, _i32, _p, _p, _i32, _i64, _i32, _p, _p, _p, _p,
                                  _p, _sz, _p]),
    "eac_cc_workspace_bytes": (_sz, [_i32, _i64]),
    "eac_connected_components": (C.c_int, [_p, _p, _p, _i32, _i64, _p, _p, _p, _sz, _p]),
    "eac_connected_components_noalloc_max_bytes": (_sz, []),
    "eac_connected_components_noalloc": (C.c_int, [_p, _p, _p, _i32, _i64, _p, _p, _p]),
    "eac_sl_lifetime_workspace_bytes": (_sz, [_i32, _i64]),
    "eac_sl_lifetime": (C.c_int, [_p, _p, _i32, _p, _p, _i32, _i64, _i32, _i32, C.c_double,
                                  _i32, _p, _p, _p, _sz, _p]),
    "eac_sl_lifetime_batched_workspace_bytes": (_sz, [_i32, _i64, _i32]),
    "eac_sl_lifetime_batched": (C.c_int, [_p, _p, _i32, _p, _p, _i32, _i64, _i32, _p, _i32, _i32,
                                          C.c_double, _i32, _p, _p, _p, _p, _sz, _p]),
    "eac_sl_lifetime_batched_graph_create": (C.c_int, [_p, _p, _i32, _p, _p, _i32, _i64, _i32,
                                                       _p, _i32, _i32, C.c_double, _i32, _i32,
                                                       _p, _p, _p, _p, _sz, _p]),
    "eac_sl_lifetime_batched_graph_launch": (C.c_int, [_p, _p]),
    "eac_sl_lifetime_batched_graph_rounds": (C.c_int, [_p]),
    "eac_mst_round_bound": (C.c_int, [_i32, _i64, _i32]),
    "eac_sl_lifetime_batched_graph_destroy": (C.c_int, [_p]),
    "eac_tune_set": (C.c_int, [_i32, _i32]),
    "eac_tune_get": (C.c_int, [_i32, _p]),
    "eac_sort_workspace_bytes": (_sz, [_i64, _i32, _i32]),
    "eac_radix_sort_u64": (C.c_int, [_p, _p, _p, _p, _i64, _i32, _i32, _p, _sz, _p]),
    "eac_kmeans_workspace_bytes": (_sz, [_i64, _i32, _i64, _i32]),
    "eac_kmeans_label": (C.c_int, [_p, _i64, _i32, _p, _p, _i32, _i32, _p, _p, _p, _p, _sz, _p]),
    "eac_kmeans_update": (C.c_int, [_p, _i64, _i32, _p, _p, _p, _i32, _p, _p, _p, _sz, _p]),
    "eac_kmeans_fit": (C.c_int, [_p, _i64, _i32, _p, _p, _i32, _i32, C.c_double, _i32, _p, _p,
                                 _p, _p, _p, _sz, _p]),
    "eac_sum_f32_workspace_bytes": (_sz, [_i64, _i64]),
    "eac_sum_f32_rows": (C.c_int, [_p, _i64, _i64, _p, _p, _sz, _p]),
    "eac_graph_from_edges_workspace_bytes": (_sz, [_i32, _i64]),
    "eac_graph_from_edges": (C.c_int, [_p, _p, _i32, _p, _p, _i32, _i64, _p, _i64, _p, _p, _p,
                                       _p, _p, _p, _sz, _p]),
    "eac_coassoc_workspace_bytes": (_sz, [_i64, _i32, _i64]),
    "eac_coassoc_count": (C.c_int, [_p, _i64, _p, _i32, _i32, _p, _p, _p, _p, _sz, _p]),
    "eac_coassoc_fill": (C.c_int, [_p, _i64, _p, _i32, _i32, _p, _i32, _p, _p, _p, _sz, _p]),
    "eac_coassoc_fill_ranked": (C.c_int, [_p, _i64, _p, _i32, _i32, _p, _p, _p, _p, _p, _sz,
                                          _p]),
    "eac_kmeans_fit_graph_workspace_bytes": (_sz, [_i64, _i32, _i64, _i32]),
    "eac_kmeans_fit_graph": (C.c_int, [_p, _i64, _i32, _p, _p, _p, _i32, _i32, C.c_double, _i32,
                                       _p, _p, _p, _p, _p, _p, _sz, _p]),
    "eac_coassoc_estimate": (C.c_int, [_p, _i64, _p, _i32, _i32, _i64, _p, _p, _sz, _p]),
    "eac_coassoc_stage_bytes": (_sz, [_i64, _i32, _i32, _i64]),
    "eac_coassoc_stage": (C.c_int, [_p, _i64, _p, _i32, _i32, _p, _p, _sz, _p, _p, _p, _sz,
                                    _p]),
    "eac_coassoc_compact": (C.c_int, [_i64, _p, _p, _sz, _i32, _p, _p, _p]),
    "eac_coassoc_graph": (C.c_int, [_p, _i64, _p, _p, _i32, _i32, _i32, _p, _p, _p, _i64, _p, _sz,
                                    _p, _p, _sz, _p]),
    "eac_coassoc_cap_workspace_bytes": (_sz, [_i64]),
    "eac_coassoc_cap": (C.c_int, [_p, _p, _p, _p, _i64, _i32, _i32, _p, _i64, _i64, _p, _p, _p,
                                  _p, _p, _sz, _p]),
    "eac_relabel_workspace_bytes": (_sz, [_i32, _i32]),
    "eac_canonical_relabel": (C.c_int, [_p, _i32, _i32, _p, _p, _p, _sz, _p]),
    "eac_contingency": (C.c_int, [_p, _p, _i64, _i32, _i32, _i32, _i32, _p, _p]),
    "eac_partition_to_bin": (C.c_int, [_p, _p, _i64, _i32, _i64, _p, _p, _p]),
    "eac_coassoc_dense": (C.c_int, [_p, _i64, _i32, _i32, _i32, _p, _p, _p, _sz, _p]),
    "eac_dense_diassoc": (C.c_int, [_p, _i64, _i32, C.c_double, _p]),
    "eac_dense_degree": (C.c_int, [_p, _i32, _i64, _i32, _p, _p, _p, _sz, _p]),
    "eac_dense_sl_workspace_bytes": (_sz, [_i64, _i32, _i32]),
    "eac_full_sl_lifetime": (C.c_int, [_p, _i32, _i64, _i32, _i32, _p, _p, _p, _p, _sz, _p]),
    "eac_slink": (C.c_int, [_p, _i32, _i64, _i32, _p, _p, _sz, _p]),
    "eac_labels_from_Z_workspace_bytes": (_sz, [_i64]),
    "eac_labels_from_Z": (C.c_int, [_p, _i64, _i32, _p, _p, _sz, _p]),
}


class EACUnavailable(RuntimeError):
    """The HIP library or a GPU is not available: there is no CPU fallback."""


class EACError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[eac {code}] {msg}")
        self.code = code


_lock = threading.Lock()
_lib = None


def load_library() -> C.CDLL:
    """Load libeac_hip.so (no GPU needed); raises EACUnavailable if absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not _LIB_PATH.exists():
            raise EACUnavailable(
                f"{_LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        # torch first: its HIP libraries ask for the runtime as libamdhip64.so
        # and this library as libamdhip64.so.7, so loaded the other way round
        # the process holds two HIP (and HSA) runtimes, and calls on the
        # second fail (a pinned host allocation, for one). Loaded after
        # torch's, the name resolves to that same runtime.
        try:
            import torch as _torch  # noqa: F401
        except ImportError:  # standalone use of the C library
            pass
        lib = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int) -> None:
    if rc == EAC_OK:
        return
    msg = load_library().eac_last_error().decode(errors="replace")
    if rc == EAC_ERR_VALUE:
        raise ValueError(msg)
    if rc in (EAC_ERR_ARG, EAC_ERR_NOT_FOUND):
        raise EACError(rc, msg)
    raise EACError(rc, msg)


# ------------------------------------------------------------ torch plumbing
def torch():
    import torch as _t  # deferred: importing torch is slow on a cold box
    return _t


def device():
    t = torch()
    if not t.cuda.is_available():
        raise EACUnavailable("no GPU visible: the EAC hot path runs only on the GPU")
    return t.device("cuda", t.cuda.current_device())


def stream_ptr():
    return C.c_void_p(torch().cuda.current_stream().cuda_stream)


_ws_cache: dict = {}


def workspace(nbytes: int):
    """Device workspace (uint8 tensor) of at least nbytes, reused per device."""
    t = torch()
    dev = device()
    key = (dev.index, t.cuda.current_stream().cuda_stream)
    cur = _ws_cache.get(key)
    if cur is None or cur.numel() < nbytes:
        cur = t.empty(max(int(nbytes), 256), dtype=t.uint8, device=dev)
        _ws_cache[key] = cur
    return cur


def ptr(x):
    return C.c_void_p(0 if x is None else x.data_ptr())


def is_torch(x) -> bool:
    try:
        import torch as _t
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, _t.Tensor)


_NP_TO_T = None


def as_device(x, dtype=None):
    """numpy array / torch tensor -> contiguous device tensor (copy only if needed)."""
    t = torch()
    global _NP_TO_T
    if _NP_TO_T is None:
        _NP_TO_T = {
            np.dtype(np.int32): t.int32, np.dtype(np.int64): t.int64,
            np.dtype(np.float32): t.float32, np.dtype(np.float64): t.float64,
            np.dtype(np.uint8): t.uint8, np.dtype(np.int16): t.int16,
            np.dtype(np.uint16): t.uint16, np.dtype(np.uint32): t.uint32,
            np.dtype(np.int8): t.int8, np.dtype(np.uint64): t.uint64,
        }
    dev = device()
    if is_torch(x):
        y = x
        if dtype is not None and y.dtype != dtype:
            y = y.to(dtype)
        if y.device != dev:
            y = y.to(dev)
        return y.contiguous()
    a = np.ascontiguousarray(x)
    y = t.from_numpy(a).to(dev)
    if dtype is not None and y.dtype != dtype:
        y = y.to(dtype)
    return y


def weight_code(dt) -> int:
    t = torch()
    table = {t.int32: W_I32, t.float32: W_F32, t.uint16: W_U16, t.uint8: W_U8,
             t.int64: W_I64, t.int16: W_I16, t.uint32: W_U32}
    if dt not in table:
        raise TypeError(f"unsupported weight dtype {dt}")
    return table[dt]
