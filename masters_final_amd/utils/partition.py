"""Ensemble generation and the partition formats (MyML/MyML/utils/partition.py).

convertClusterStringToIndex (partition.py:73-86) and generateEnsemble
(partition.py:88-124) keep the reference's signatures. The grouping of
samples by label runs on the GPU (a stable radix sort of (label, index) in
libeac_hip.so); the split into per-cluster arrays is a host view of that
order. generate_ensemble_batched runs all partitions' K-Means in one batched
call instead of one fit per partition.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib as L


def _group_by_label(labels):
    """Stable group-by on the GPU: (order int64 [n], counts per value)."""
    t = L.torch()
    lib = L.load_library()
    d = L.as_device(labels, t.int64)
    n = d.numel()
    lo = int(d.min().item()) if n else 0
    shifted = (d - lo).to(t.int64)
    span = int(shifted.max().item()) + 1 if n else 1
    keys = (shifted << 32) | t.arange(n, device=d.device, dtype=t.int64)
    out = t.empty_like(keys)
    ws = L.workspace(lib.eac_sort_workspace_bytes(n, 8, 0))
    bits = 32 + max(1, int(np.ceil(np.log2(span + 1))))
    L.check(lib.eac_radix_sort_u64(L.ptr(keys), L.ptr(out), None, None, n, 32, min(bits, 64),
                                   L.ptr(ws), ws.numel(), L.stream_ptr()))
    order = (out & 0xffffffff).cpu().numpy()
    vals = (out >> 32).cpu().numpy() + lo
    return order, vals


def _to_bin_device(rows, cols, k, N):
    """k x N uint8 device matrix with ones at (rows[t], cols[t]) (cols None:
    column t). Raises IndexError, as numpy's fancy assignment does, when an
    entry falls outside the matrix."""
    t = L.torch()
    lib = L.load_library()
    d_rows = L.as_device(rows, t.int32)
    d_cols = None if cols is None else L.as_device(cols, t.int64)
    m = d_rows.numel()
    out = t.empty((k, N), dtype=t.uint8, device=d_rows.device if m else L.device())
    bad = t.zeros(1, dtype=t.int64, device=out.device)
    L.check(lib.eac_partition_to_bin(L.ptr(d_rows) if m else None, L.ptr(d_cols), m, k, N,
                                     L.ptr(out), L.ptr(bad), L.stream_ptr()))
    nbad = int(bad.item())
    if nbad:
        raise IndexError(f"{nbad} cluster/sample indices outside the {k} x {N} matrix")
    return out


def index_to_bin_device(clusts):
    """convertIndexToBin on the GPU, result left on the device."""
    if clusts is None:
        raise Exception("A clustering partition must be provided.")
    sizes = [int(np.asarray(c).size) for c in clusts]
    N, k = sum(sizes), len(clusts)
    if N == 0:
        return _to_bin_device(np.zeros(0, np.int32), np.zeros(0, np.int64), k, 0)
    rows = np.repeat(np.arange(k, dtype=np.int32), sizes)
    cols = np.concatenate([np.asarray(c).astype(np.int64).reshape(-1) for c in clusts])
    return _to_bin_device(rows, cols, k, N)


def convertIndexToBin(clusts):
    """partition.py:17-35: list of index arrays -> k x N uint8 binary matrix
    (N = number of listed indices), built on the GPU."""
    return index_to_bin_device(clusts).cpu().numpy()


def string_to_bin_device(clusts):
    """convertClusterStringToBin on the GPU, result left on the device."""
    if clusts is None:
        raise Exception("A clustering partition must be provided.")
    t = L.torch()
    d = L.as_device(np.asarray(clusts).astype(np.int64) if not L.is_torch(clusts) else clusts,
                    t.int64)
    u = t.unique(d)
    mx, mn = int(d.max().item()), int(d.min().item())
    if u.numel() not in (mx, mx + 1):
        raise ValueError('cluster numbers go from [0, biggest_clust[ or'
                         '[1, biggest_clust[ uninterrupted')
    offset = -1 if mn == 1 else 0
    return _to_bin_device((d + offset).to(t.int32), None, u.numel(), d.numel())


def convertClusterStringToBin(clusts):
    """partition.py:37-70: label array -> k x N uint8 binary matrix (rows are
    labels, shifted down by one when the smallest label is 1), built on the
    GPU. A gapped numbering the reference's check lets through (e.g. {0, 2})
    raises IndexError, as the reference's assignment does."""
    return string_to_bin_device(clusts).cpu().numpy()


def convertClusterStringToIndex(partition):
    """partition.py:73-86: label array -> list of uint64 index arrays, one per
    distinct label in ascending label order, members ascending."""
    order, vals = _group_by_label(np.asarray(partition).astype(np.int64))
    if order.size == 0:
        return []
    cuts = np.nonzero(np.diff(vals))[0] + 1
    return [a.astype(np.uint64) for a in np.split(order, cuts)]


def generateEnsemble(data, generator, n_clusters=20, npartitions=30, iters=3):
    """partition.py:88-124: k = np.random.randint(kmin, kmax) (upper bound
    exclusive) per partition, generator.fit, conversion to index lists."""
    ensemble = [None] * npartitions
    if type(n_clusters) is list:
        if n_clusters[0] == n_clusters[1]:
            cluster_range = False
            generator.n_clusters = n_clusters[0]
        else:
            cluster_range = True
            kmin, kmax = n_clusters
    else:
        cluster_range = False
        generator.n_clusters = n_clusters
    generator.max_iter = iters
    for x in range(npartitions):
        if cluster_range:
            generator.n_clusters = np.random.randint(kmin, kmax)
        generator.fit(data)
        ensemble[x] = convertClusterStringToIndex(generator.labels_)
    return ensemble


def generate_ensemble_batched(data, k_list, iters=3, seed=0, precision="bf16x3", init=None):
    """All partitions in one batched K-Means call. Initial centroids are k
    distinct random samples per partition (numpy PCG64 seeded with `seed`)
    unless `init` (sum k x d) is given. Returns (labels [P, n] device int32,
    fit result dict)."""
    from ..cluster.KMeans import fit_batched
    X = data
    n = X.shape[0]
    if init is None:
        # the sample indices come from the host generator; the rows are
        # gathered where X lives (a device gather avoids copying X back)
        rng = np.random.default_rng(seed)
        idx = np.concatenate([rng.choice(n, int(k), replace=False) for k in k_list])
        if L.is_torch(X):
            t = L.torch()
            init = X.index_select(0, t.from_numpy(idx).to(X.device))
        else:
            init = np.asarray(X)[idx]
    r = fit_batched(X, init, list(k_list), max_iter=iters, tol=1e-4, precision=precision)
    return r["labels"], r


def n_samples_from_partition(partition):
    """partition.py:466-467."""
    return sum(map(len, partition))


def biggest_cluster_size(ensemble):
    """partition.py:469-470."""
    return max([max(map(np.size, p)) for p in ensemble])
