"""Boruvka minimum spanning forest on the GPU, behind the reference's API.

Replaces MyML/MyML/graph/mst.py:143-232 boruvka_minho_seq and mst.py:417-564
boruvka_minho_gpu. Both names run the HIP kernels of libeac_hip.so (there is
no CPU path in this package); `_seq` keeps the host-array contract of the
reference's CPU function, `_gpu` additionally accepts device tensors and can
return them (returnDevAry).

Tie-break: the reference CPU path orders a vertex's edges by (weight,
contracted destination id, CSR position) (mst.py:268-279) and its GPU path
depends on atomic order (mst.py:805), so under equal weights both may return
different trees of the same weight (test_boruvka.py:95-97). Here edges are
ordered by (weight, min(u,v), max(u,v)) on the original ids: the forest is
unique, its total weight equals the reference's, and each forest edge is
reported by the CSR position of its entry in the smaller endpoint's row.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib as L


def _graph_to_device(dest, weight, firstEdge, outDegree):
    t = L.torch()
    d_dest = L.as_device(dest, t.int32)
    d_w = L.as_device(weight)
    if d_w.dtype == t.float64:
        raise TypeError("float64 weights are not supported; pass float32 or an integer type")
    if d_w.dtype == t.int8:
        d_w = d_w.to(t.int16)
    d_fe = L.as_device(firstEdge, t.int64)
    d_od = L.as_device(outDegree, t.int32) if outDegree is not None else None
    if d_dest.numel() != d_w.numel():
        raise ValueError("dest and weight must have the same length")
    if d_od is not None and d_od.numel() != d_fe.numel():
        raise ValueError("firstEdge and outDegree must have the same length")
    return d_dest, d_w, d_fe, d_od


def boruvka_mst(dest, weight, firstEdge, outDegree=None, symmetric=True, want_ids=True,
                want_weights=False, want_labels=False):
    """Full-control entry: returns dict(ids, weights, labels, n_mst, n_components, rounds)
    with device tensors (first n_mst entries valid)."""
    t = L.torch()
    lib = L.load_library()
    d_dest, d_w, d_fe, d_od = _graph_to_device(dest, weight, firstEdge, outDegree)
    n = d_fe.numel() if d_od is not None else d_fe.numel() - 1
    m = d_dest.numel()
    dev = d_dest.device
    nf = max(n - 1, 0)
    ids = t.empty(nf, dtype=t.int64, device=dev) if want_ids else None
    wts = t.empty(nf, dtype=d_w.dtype, device=dev) if want_weights else None
    labels = t.empty(n, dtype=t.int32, device=dev) if want_labels else None
    h_out = (C.c_int64 * 4)()
    ws_bytes = lib.eac_boruvka_workspace_bytes(n, m)
    ws = L.workspace(ws_bytes)
    L.check(lib.eac_boruvka_mst(L.ptr(d_dest), L.ptr(d_w), L.weight_code(d_w.dtype), L.ptr(d_fe),
                                L.ptr(d_od), n, m, 1 if symmetric else 0, L.ptr(ids), L.ptr(wts),
                                L.ptr(labels), h_out, L.ptr(ws), ws.numel(), L.stream_ptr()))
    return {"ids": ids, "weights": wts, "labels": labels, "n_mst": int(h_out[0]),
            "n_components": int(h_out[1]), "rounds": int(h_out[2]), "weight_bits": int(h_out[3])}


def boruvka_minho_gpu(dest_in, weight_in, firstEdge_in, outDegree_in, MAX_TPB=512, stream=None,
                      returnDevAry=False):
    """mst.py:417-564. Returns (mst, n_mst): host int32 array of size n-1 whose first
    n_mst entries are forest edge ids, or device tensors (mst, n_mst[1]) when
    returnDevAry. MAX_TPB and stream are accepted for compatibility (launch
    shapes are chosen by the library; work runs on torch's current stream)."""
    r = boruvka_mst(dest_in, weight_in, firstEdge_in, outDegree_in)
    ids = r["ids"]
    if returnDevAry:
        t = L.torch()
        return ids.to(t.int32), t.tensor([r["n_mst"]], dtype=t.int32, device=ids.device)
    out = np.zeros(ids.numel(), dtype=np.int32)
    out[:] = ids.cpu().numpy()
    return out, r["n_mst"]


def boruvka_minho_seq(dest_in, weight_in, firstEdge_in, outDegree_in):
    """mst.py:143-232 signature (host arrays in, (mst, n_mst) out), computed on the GPU."""
    return boruvka_minho_gpu(dest_in, weight_in, firstEdge_in, outDegree_in)


def compute_cuda_grid_dim(n, tpb):
    """mst.py:568-571 helper kept for callers that size their own launches."""
    return int(np.ceil(n / np.float32(tpb)))
