"""Connected components on the GPU, behind the reference's API.

Replaces MyML/MyML/graph/connected_components.py:46-137 connected_comps_seq
and :176-314 connected_comps_gpu. The reference contracts the graph with
Boruvka and returns final supervertex ids; this returns the same partition
numbered in order of first occurrence (vertex 0 in component 0), computed by
a lock-free union-find over the CSR entries (weights are not needed).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib as L


def connected_components_device(dest, firstEdge, outDegree=None):
    """Device labels (int32 tensor) and the number of components."""
    t = L.torch()
    lib = L.load_library()
    d_dest = L.as_device(dest, t.int32)
    d_fe = L.as_device(firstEdge, t.int64)
    d_od = L.as_device(outDegree, t.int32) if outDegree is not None else None
    n = d_fe.numel() if d_od is not None else d_fe.numel() - 1
    m = d_dest.numel()
    labels = t.empty(n, dtype=t.int32, device=d_dest.device)
    nc = C.c_int64()
    ws = L.workspace(lib.eac_cc_workspace_bytes(n, m))
    L.check(lib.eac_connected_components(L.ptr(d_dest), L.ptr(d_fe), L.ptr(d_od), n, m,
                                         L.ptr(labels), C.byref(nc), L.ptr(ws), ws.numel(),
                                         L.stream_ptr()))
    return labels, int(nc.value)


def connected_comps_gpu(dest_in, weight_in, firstEdge_in, outDegree_in, MAX_TPB=512,
                        stream=None):
    """connected_components.py:176-314: host int32 labels."""
    labels, _ = connected_components_device(dest_in, firstEdge_in, outDegree_in)
    return labels.cpu().numpy()


def connected_comps_seq(dest_in, weight_in, firstEdge_in, outDegree_in):
    """connected_components.py:46-137 signature, computed on the GPU."""
    return connected_comps_gpu(dest_in, weight_in, firstEdge_in, outDegree_in)


def canonical_relabel(labels, n_values=None):
    """Renumber labels by first occurrence on the GPU; returns (tensor, count)."""
    t = L.torch()
    lib = L.load_library()
    d = L.as_device(labels, t.int32)
    n = d.numel()
    if n_values is None:
        n_values = int(d.max().item()) + 1 if n else 0
    out = t.empty(n, dtype=t.int32, device=d.device)
    cnt = C.c_int64()
    ws = L.workspace(lib.eac_relabel_workspace_bytes(n, n_values))
    L.check(lib.eac_canonical_relabel(L.ptr(d), n, n_values, L.ptr(out), C.byref(cnt),
                                      L.ptr(ws), ws.numel(), L.stream_ptr()))
    return out, int(cnt.value)
