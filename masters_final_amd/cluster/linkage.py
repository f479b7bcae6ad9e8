"""Single-linkage helpers on the GPU behind the reference's API
(MyML/MyML/cluster/linkage.py, imported by the EAC driver as SL and
clusters_from_SL).

scipy_numba_slink_wraper (linkage.py:144-149, SL) returns the float32
linkage matrix Z of single linkage on condensed (or square) dissimilarities;
the rows come from the minimum spanning tree of the complete graph, built by
Borůvka in libeac_hip.so, in ascending height, rows of equal height in
(height, min id, max id) order of their tree edges (the reference's SLINK
orders equal heights by its own quicksort; the heights, the sizes and every
cut between distinct heights are the same). labels_from_Z_numba
(linkage.py:303-357, clusters_from_SL, the function the EAC driver calls)
applies the first n - n_clusters merges of any valid Z on the GPU.
"""
from __future__ import annotations

import numpy as np

from .. import _lib as L


def slink_device(weights, n):
    """Z ((n-1) x 4 float32 device tensor) of single linkage."""
    from ..EAC.full import _wcode
    t = L.torch()
    lib = L.load_library()
    d = L.as_device(weights)
    if d.dtype == t.float64:
        d = d.to(t.float32)
    condensed = d.dim() == 1
    n = int(n)
    code = _wcode(d)
    Z = t.empty((max(n - 1, 1), 4), dtype=t.float32, device=d.device)
    ws = L.workspace(lib.eac_dense_sl_workspace_bytes(n, code, 1 if condensed else 0))
    L.check(lib.eac_slink(L.ptr(d), code, n, 1 if condensed else 0, L.ptr(Z), L.ptr(ws),
                          ws.numel(), L.stream_ptr()))
    return Z[: n - 1]


def scipy_numba_slink_wraper(weights, n):
    """linkage.py:144-149: Z = single linkage of the condensed weights
    (float32, host)."""
    return slink_device(weights, n).cpu().numpy()


SL = scipy_numba_slink_wraper


def labels_from_Z_device(Z, n_clusters):
    """First n - n_clusters merges of Z; labels [n] int32 (device) numbered
    by first occurrence."""
    t = L.torch()
    lib = L.load_library()
    dZ = L.as_device(Z, t.float32).reshape(-1, 4)
    n = dZ.shape[0] + 1
    labels = t.empty(n, dtype=t.int32, device=dZ.device)
    ws = L.workspace(lib.eac_labels_from_Z_workspace_bytes(n))
    L.check(lib.eac_labels_from_Z(L.ptr(dZ), n, int(n_clusters), L.ptr(labels), L.ptr(ws),
                                  ws.numel(), L.stream_ptr()))
    return labels


def labels_from_Z_numba(Z, track, n_clusters):
    """linkage.py:303-357: fills track (size n) with the labels of the first
    n - n_clusters merges, numbered by first occurrence."""
    lab = labels_from_Z_device(Z, n_clusters)
    if L.is_torch(track):
        track.copy_(lab.to(track.dtype))
    else:
        track[...] = lab.cpu().numpy().astype(np.asarray(track).dtype)


clusters_from_SL = labels_from_Z_numba
