"""Exclusive prefix sums on the GPU, behind the reference's API.

Replaces MyML/MyML/utils/scan.py:144-249 scan_gpu (and the CPU twins
scan.py:41-53 exprefixsumNumba, scan.py:56-72 exprefixsumNumbaSingle for the
integer arrays the EAC path scans). Integer dtypes only, as the reference's
GPU scan (scan.py:170-171): a parallel float scan cannot reproduce the
sequential rounding of the CPU loop, so float input raises TypeError.
"""
from __future__ import annotations

import numpy as np

from .. import _lib as L


def _scan_device(d, init=0):
    t = L.torch()
    lib = L.load_library()
    n = d.numel()
    total = t.empty(1, dtype=d.dtype, device=d.device)
    if d.dtype == t.int32:
        ws = L.workspace(lib.eac_scan_workspace_bytes(n, 4))
        L.check(lib.eac_scan_exclusive_i32(L.ptr(d), L.ptr(d), n, int(init), L.ptr(total),
                                           L.ptr(ws), ws.numel(), L.stream_ptr()))
    elif d.dtype == t.int64:
        ws = L.workspace(lib.eac_scan_workspace_bytes(n, 8))
        L.check(lib.eac_scan_exclusive_i64(L.ptr(d), L.ptr(d), n, int(init), L.ptr(total),
                                           L.ptr(ws), ws.numel(), L.stream_ptr()))
    else:
        raise TypeError(f"scan: integer arrays only (got {d.dtype})")
    return total


def _as_int_device(a):
    t = L.torch()
    if L.is_torch(a):
        if a.dtype not in (t.int32, t.int64):
            raise TypeError(f"scan: integer arrays only (got {a.dtype})")
        return L.as_device(a)
    a = np.asarray(a)
    if not np.issubdtype(a.dtype, np.integer):
        raise TypeError(f"scan: integer arrays only (got {a.dtype})")
    return L.as_device(a, t.int64 if a.dtype.itemsize > 4 or a.dtype == np.uint32 else t.int32)


def scan_gpu(in_ary, MAX_TPB=512, stream=0):
    """In-place exclusive scan. Device tensor in: returns the device total (shape (1,)).
    numpy in: the scanned values are written back into the array (the
    reference discards them, scan.py:164-165) and the total is returned as a
    host scalar."""
    if L.is_torch(in_ary) and in_ary.is_cuda:
        if not in_ary.is_contiguous():
            raise ValueError("scan_gpu: device array must be contiguous")
        return _scan_device(_as_int_device(in_ary))
    a = np.asarray(in_ary)
    d = _as_int_device(a)
    total = _scan_device(d)
    a[...] = d.cpu().numpy().astype(a.dtype, copy=False).reshape(a.shape)
    return total.cpu().numpy()[0].item()


def exprefixsumNumba(in_ary, out_ary, init=0):
    """scan.py:41-53: out_ary <- exclusive scan of in_ary offset by init; returns init+sum."""
    d = _as_int_device(in_ary).clone()
    total = _scan_device(d, init)
    if L.is_torch(out_ary):
        out_ary.copy_(d.to(out_ary.dtype))
    else:
        out_ary[...] = d.cpu().numpy().astype(np.asarray(out_ary).dtype, copy=False)
    return total.cpu().numpy()[0].item()


def exprefixsumNumbaSingle(in_ary, init=0):
    """scan.py:56-72: in-place exclusive scan offset by init; returns init+sum."""
    if L.is_torch(in_ary) and in_ary.is_cuda:
        d = _as_int_device(in_ary)
        total = _scan_device(d, init)
        if d.data_ptr() != in_ary.data_ptr():
            in_ary.copy_(d)
        return total.cpu().numpy()[0].item()
    a = np.asarray(in_ary)
    d = _as_int_device(a)
    total = _scan_device(d, init)
    a[...] = d.cpu().numpy().astype(a.dtype, copy=False)
    return total.cpu().numpy()[0].item()
