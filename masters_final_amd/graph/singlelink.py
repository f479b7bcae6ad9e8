"""Single-link clustering by a largest-lifetime cut of the minimum spanning
forest, on the GPU, behind the reference's API.

Replaces MyML/MyML/graph/singlelink.py:14-65 sl_mst_lifetime_seq and the
intended GPU twin sl_mst_lifetime_gpu (singlelink.py:69-153, which cannot be
imported: SyntaxError at :4-8 and undefined names at :134, :186). The forest,
the sort of its weights, the lifetime rule, the cut and the components all
run in libeac_hip.so (eac_sl_lifetime, rule EAC_CUT_SINGLELINK).

Semantics kept from the reference: disconnect_weight defaults to the largest
graph weight; the forest keeps its a smallest edges (a = first index of the
largest lifetime) when that lifetime is >= disconnect_weight - largest forest
weight, i.e. the edge at the lower end of the gap is dropped too
(singlelink.py:43-48); fewer than two forest edges raise ValueError (numpy's
argmax of an empty array). Differences: ties among equal weights follow the
(weight, min id, max id) order instead of numpy's unstable argsort, and the
labels are numbered by first occurrence instead of Boruvka supervertex ids.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib as L
from .mst import _graph_to_device

EAC_CUT_EAC = 0
EAC_CUT_SINGLELINK = 1


def lifetime_cut(dest, weight, firstEdge, outDegree=None, rule=EAC_CUT_SINGLELINK,
                 max_dist=None, n_clusters=0, symmetric=True):
    """Device labels (int32 tensor, or None when the rule yields one cluster
    under EAC_CUT_EAC) and a dict of the cut's counters."""
    t = L.torch()
    lib = L.load_library()
    d_dest, d_w, d_fe, d_od = _graph_to_device(dest, weight, firstEdge, outDegree)
    n = d_fe.numel() if d_od is not None else d_fe.numel() - 1
    m = d_dest.numel()
    labels = t.empty(max(n, 1), dtype=t.int32, device=d_dest.device)
    h = (C.c_int64 * 6)()
    ws = L.workspace(lib.eac_sl_lifetime_workspace_bytes(n, m))
    L.check(lib.eac_sl_lifetime(L.ptr(d_dest), L.ptr(d_w), L.weight_code(d_w.dtype), L.ptr(d_fe),
                                L.ptr(d_od), n, m, 1 if symmetric else 0, int(rule),
                                -1.0 if max_dist is None else float(max_dist), int(n_clusters),
                                L.ptr(labels), h, L.ptr(ws), ws.numel(), L.stream_ptr()))
    info = {"n_components": int(h[0]), "labels_empty": bool(h[1]), "nc_stable": int(h[2]),
            "kept_edges": int(h[3]), "forest_edges": int(h[4]), "rounds": int(h[5])}
    return (None if info["labels_empty"] else labels[:n]), info


def sl_mst_lifetime_gpu(dest, weight, fe, od, disconnect_weight=None, MAX_TPB=256,
                        stream=None):
    """singlelink.py:69-153 signature: host int32 labels."""
    labels, _ = lifetime_cut(dest, weight, fe, od, rule=EAC_CUT_SINGLELINK,
                             max_dist=disconnect_weight)
    return labels.cpu().numpy()


def sl_mst_lifetime_seq(dest, weight, fe, od, disconnect_weight=None):
    """singlelink.py:14-65 signature (host arrays), computed on the GPU."""
    return sl_mst_lifetime_gpu(dest, weight, fe, od, disconnect_weight)


def lifetime_cut_batched(dest, weight, indptr, goff, rule=EAC_CUT_EAC, max_dist=None,
                         n_clusters=0, symmetric=True):
    """Many independent graphs as one disjoint union (graph g owns vertices
    goff[g]:goff[g+1] and their rows; dest holds union ids). Returns (labels
    per graph, numbered by first occurrence inside each graph, as one device
    int32 tensor [N]; per-graph int64 tensor [G, 4] = components,
    labels_empty, nc_stable, kept edges (-1 where the reference raises))."""
    t = L.torch()
    lib = L.load_library()
    d_dest, d_w, d_ip, _ = _graph_to_device(dest, weight, indptr, None)
    N = d_ip.numel() - 1
    M = d_dest.numel()
    g = np.ascontiguousarray(np.asarray(goff, dtype=np.int64))
    G = g.size - 1
    labels = t.empty(N, dtype=t.int32, device=d_dest.device)
    gout = t.empty((G, 4), dtype=t.int64, device=d_dest.device)
    h = (C.c_int64 * 4)()
    ws = L.workspace(lib.eac_sl_lifetime_batched_workspace_bytes(N, M, G))
    L.check(lib.eac_sl_lifetime_batched(L.ptr(d_dest), L.ptr(d_w), L.weight_code(d_w.dtype),
                                        L.ptr(d_ip), None, N, M, 1 if symmetric else 0,
                                        g.ctypes.data_as(C.c_void_p), G, int(rule),
                                        -1.0 if max_dist is None else float(max_dist),
                                        int(n_clusters), L.ptr(labels), L.ptr(gout), h,
                                        L.ptr(ws), ws.numel(), L.stream_ptr()))
    return labels, gout, {"components": int(h[0]), "forest_edges": int(h[1]),
                          "rounds": int(h[2]), "kept_edges": int(h[3])}
