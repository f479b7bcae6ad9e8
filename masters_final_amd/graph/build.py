"""Graph rebuilt from selected edges on the GPU, behind the reference's API
(MyML/MyML/graph/build.py).

getGraphFromEdges_seq (build.py:49-93) fills the caller's arrays and returns
0, or -1 when an edge id has no origin vertex; getGraphFromEdges_gpu
(build.py:96-151) returns (ndest, nweight, nfe, nod) device tensors and skips
ids of -1. Both run eac_graph_from_edges; rows keep the CPU function's list
order (the reference's GPU version leaves it to atomic order).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib as L


def _gfe(dest, weight, fe, od, edges, n_edges=None):
    t = L.torch()
    lib = L.load_library()
    d_dest = L.as_device(dest, t.int32)
    d_w = L.as_device(weight)
    d_fe = L.as_device(fe, t.int64)
    d_od = L.as_device(od, t.int32) if od is not None else None
    d_e = L.as_device(edges, t.int64)
    if n_edges is not None:
        k = int(n_edges.item() if L.is_torch(n_edges) else np.asarray(n_edges).reshape(-1)[0])
        d_e = d_e[:k]
    n = d_fe.numel() if d_od is not None else d_fe.numel() - 1
    k = d_e.numel()
    nod = t.empty(n, dtype=t.int32, device=d_dest.device)
    nfe = t.empty(n, dtype=t.int64, device=d_dest.device)
    ndest = t.empty(max(2 * k, 1), dtype=t.int32, device=d_dest.device)
    nw = t.empty(max(2 * k, 1), dtype=d_w.dtype, device=d_dest.device)
    nitems = C.c_int64()
    ws = L.workspace(lib.eac_graph_from_edges_workspace_bytes(n, k))
    rc = lib.eac_graph_from_edges(L.ptr(d_dest), L.ptr(d_w), d_w.element_size(), L.ptr(d_fe),
                                  L.ptr(d_od), n, d_dest.numel(), L.ptr(d_e), k, L.ptr(nod),
                                  L.ptr(nfe), L.ptr(ndest), L.ptr(nw), C.byref(nitems), L.ptr(ws),
                                  ws.numel(), L.stream_ptr())
    return rc, nod, nfe, ndest[: nitems.value], nw[: nitems.value]


def getGraphFromEdges_seq(dest, weight, fe, od, edges, nod, nfe, ndest, nweight):
    """build.py:49-93: fills nod, nfe, ndest, nweight; returns 0 or -1."""
    rc, d_nod, d_nfe, d_nd, d_nw = _gfe(dest, weight, fe, od, edges)
    if rc == L.EAC_ERR_NOT_FOUND:
        return -1
    L.check(rc)
    for host, dev in ((nod, d_nod), (nfe, d_nfe), (ndest, d_nd), (nweight, d_nw)):
        if L.is_torch(host):
            host[: dev.numel()].copy_(dev.to(host.dtype))
        else:
            host[: dev.numel()] = dev.cpu().numpy().astype(np.asarray(host).dtype, copy=False)
    return 0


def getGraphFromEdges_gpu(dest, weight, fe, od, edges, n_edges=None, MAX_TPB=512, stream=None):
    """build.py:96-151: (ndest, nweight, nfe, nod) device tensors."""
    rc, nod, nfe, nd, nw = _gfe(dest, weight, fe, od, edges, n_edges)
    L.check(rc)
    return nd, nw, nfe, nod


def outdegree_from_firstedge(firstedge, outdegree, n_edges):
    """build.py:275-280: outdegree[v] = firstedge[v+1] - firstedge[v], last uses n_edges."""
    t = L.torch()
    d = L.as_device(firstedge, t.int64)
    od = t.empty_like(d)
    od[:-1] = d[1:] - d[:-1]
    od[-1] = int(n_edges) - d[-1]
    if L.is_torch(outdegree):
        outdegree.copy_(od.to(outdegree.dtype))
    else:
        outdegree[...] = od.cpu().numpy().astype(np.asarray(outdegree).dtype)
    return outdegree
