"""Contingency table of two partitions, counted by `eac_contingency`.

The reference builds a binary cluster-by-sample matrix per partition
(MyML/MyML/utils/partition.py:17-70) and multiplies them
(MyML/MyML/validation/consistency.py:87-101, hindex.py:44). For partitions the
product is the contingency table T[i][j] = #{s : sample s in cluster i of the
first partition and in cluster j of the second}, which the GPU counts in one
pass over the two label arrays.

Partition formats (the reference's `format` argument):
  'array' -- length-N labels numbered [0, K) or [1, K] without gaps (any other
             numbering raises ValueError, partition.py:56-59); row i is label
             i (0-based) or i + 1 (1-based);
  'list'  -- list of index arrays, one per cluster, in order; samples in no
             cluster are counted in an extra last row/column, so cluster sizes
             still equal the arrays' sizes;
  'bin'   -- binary cluster-by-sample matrix whose columns hold at most one 1
             (a partition); converted to 'list'.
"""
from __future__ import annotations

import numpy as np

from .. import _lib as L


def _labels_from_list(clusts, N=None):
    sizes = [int(np.asarray(c).size) for c in clusts]
    n = int(N) if N is not None else int(sum(sizes))
    k = len(clusts)
    lab = np.full(n, k, dtype=np.int32)  # k = "in no cluster"
    seen = np.zeros(n, dtype=np.int64)
    for i, c in enumerate(clusts):
        idx = np.asarray(c, dtype=np.int64)
        if idx.size and (idx.min() < 0 or idx.max() >= n):
            raise IndexError("cluster index out of range")
        np.add.at(seen, idx, 1)
        lab[idx] = i
    if np.any(seen > 1):
        # a partition: the 'bin' format rejects the same overlap
        raise ValueError("list format: a sample is listed in more than one cluster")
    return lab, k, np.asarray(sizes, dtype=np.int64)


def _labels_from_bin(mat):
    m = np.asarray(mat)
    if m.ndim != 2:
        raise TypeError("Array must have two dimensions")
    nz = m != 0
    per_col = nz.sum(axis=0)
    if np.any(per_col > 1):
        raise ValueError("bin format: a sample belongs to more than one cluster")
    return [np.nonzero(nz[i])[0] for i in range(m.shape[0])]


def _as_rows(clusts, fmt, N=None):
    """-> (device int32 labels, first label, number of rows, has 'none' row,
    cluster sizes or None)."""
    t = L.torch()
    if fmt == "bin":
        width = int(np.asarray(clusts).shape[1]) if np.ndim(clusts) == 2 else None
        if N is not None and width is not None and int(N) != width:
            raise ValueError(f"bin format: N = {N} but the matrix has {width} columns")
        clusts, fmt = _labels_from_bin(clusts), "list"
        N = width  # every column is a sample, with or without a cluster
    if fmt == "list":
        lab, k, sizes = _labels_from_list(clusts, N)
        return L.as_device(lab, t.int32), 0, k + 1, True, sizes
    if fmt != "array":
        raise ValueError("Format not accepted: {}".format(fmt))
    d = L.as_device(clusts, t.int64)
    if d.numel() == 0:
        raise ValueError("empty partition")
    return d, None, None, False, None  # range read back by contingency()


def contingency(clusts1, clusts2, format1="array", format2="array", N=None):
    """Host int64 table T [rows1, rows2] (rows in cluster order, plus a last
    'none' row/column for 'list' inputs) and the two partitions' cluster sizes."""
    t = L.torch()
    lib = L.load_library()
    a, a0, ka, none1, size1 = _as_rows(clusts1, format1, N)
    b, b0, kb, none2, size2 = _as_rows(clusts2, format2, N)
    if a.numel() != b.numel():
        raise ValueError("the two partitions label different numbers of samples")
    # label ranges of 'array' inputs: one read-back for both
    rows = [[a, a0, ka], [b, b0, kb]]
    ext = [q for q in range(2) if rows[q][2] is None]
    if ext:
        mm = t.stack([v for q in ext for v in (rows[q][0].min(), rows[q][0].max())])
        mm = mm.cpu().tolist()
        for e, q in enumerate(ext):
            lo, hi = int(mm[2 * e]), int(mm[2 * e + 1])
            # partition.py:56-59: labels [0, K) or [1, K] without gaps; the
            # distinct count is checked before any table is allocated
            if lo not in (0, 1) or hi >= 2 ** 31 - 1 or \
                    int(t.unique(rows[q][0]).numel()) != hi - lo + 1:
                raise ValueError('cluster numbers go from [0, biggest_clust[ or'
                                 '[1, biggest_clust[ uninterrupted')
            rows[q] = [rows[q][0].to(t.int32).contiguous(), lo, hi - lo + 1]
    (a, a0, ka), (b, b0, kb) = rows
    if ka * kb > 2 ** 31:
        raise ValueError(f"contingency table of {ka} x {kb} cells is over the 2^31 limit")
    out = t.empty(ka * kb + 1, dtype=t.int64, device=a.device)
    L.check(lib.eac_contingency(L.ptr(a), L.ptr(b), a.numel(), a0, ka, b0, kb, L.ptr(out),
                                L.stream_ptr()))
    h = out.cpu().numpy()
    if h[-1]:
        raise ValueError("labels outside the partition's cluster range")
    T = h[:-1].reshape(ka, kb)
    s1 = T.sum(axis=1) if size1 is None else size1
    s2 = T.sum(axis=0) if size2 is None else size2
    if none1:
        T = T[:-1]
    if none2:
        T = T[:, :-1]
    return T, np.asarray(s1, np.int64), np.asarray(s2, np.int64), int(a.numel())
