"""Dense co-association matrix and its single-link cut on the GPU, behind the
reference's API (MyML/MyML/EAC/full.py, MyML/MyML/EAC/eac.py:337-391).

EAC_FULL keeps the reference's constructor and update_ensemble / get_degree /
make_diassoc / todense / tocsr sequence (full.py:8-131). The matrix lives on
the device as a torch tensor in the requested dtype (uint8, uint16, int32 or
float32; the reference's uint8 default wraps at 256 like its += 1 does):
condensed (n(n-1)/2, SciPy's order) or n x n. Counting, dissimilarities and
degrees run in libeac_hip.so; full_sl_lifetime runs single linkage (the
minimum spanning tree of the complete graph), the lifetime rule and the
labels there too. It is the reference's O(n^2) cross-check of the sparse
path for small n (n <= 65536 here).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib as L

_DT = {np.dtype(np.uint8): L.W_U8, np.dtype(np.uint16): L.W_U16,
       np.dtype(np.int32): L.W_I32, np.dtype(np.float32): L.W_F32}


def _wcode(tensor):
    t = L.torch()
    code = {t.uint8: L.W_U8, t.uint16: L.W_U16, t.int32: L.W_I32, t.float32: L.W_F32}.get(
        tensor.dtype)
    if code is None:
        raise TypeError(f"dense matrices of {tensor.dtype} are not supported "
                        "(uint8, uint16, int32, float32)")
    return code


def coassoc_dense_device(labels, condensed=True, dtype=np.uint8):
    """EAC_FULL counts of labels [P, n] on the GPU: (matrix tensor, largest
    stored value)."""
    t = L.torch()
    lib = L.load_library()
    dl = L.as_device(labels, t.int32)
    if dl.dim() == 1:
        dl = dl[None]
    P, n = dl.shape
    dt = np.dtype(dtype)
    if dt not in _DT:
        raise TypeError("dtype must be uint8, uint16, int32 or float32")
    tdt = {L.W_U8: t.uint8, L.W_U16: t.uint16, L.W_I32: t.int32, L.W_F32: t.float32}[_DT[dt]]
    shape = (n * (n - 1) // 2,) if condensed else (n, n)
    out = t.empty(shape, dtype=tdt, device=dl.device)
    mx = C.c_int64()
    ws = L.workspace(256)
    L.check(lib.eac_coassoc_dense(L.ptr(dl), n, P, 1 if condensed else 0, _DT[dt], L.ptr(out),
                                  C.byref(mx), L.ptr(ws), ws.numel(), L.stream_ptr()))
    return out, mx.value


def full_sl_lifetime_device(mat, n_samples, n_clusters=0, want_Z=False):
    """eac.py:337-364 on a device dissimilarity matrix (condensed or n x n):
    (n_clusters, labels tensor or None for one cluster, Z tensor or None)."""
    t = L.torch()
    lib = L.load_library()
    d = L.as_device(mat)
    condensed = d.dim() == 1
    n = int(n_samples)
    if (condensed and d.numel() != n * (n - 1) // 2) or (not condensed and d.shape != (n, n)):
        raise ValueError("matrix shape does not match n_samples")
    code = _wcode(d)
    labels = t.empty(n, dtype=t.int32, device=d.device)
    Z = t.empty((max(n - 1, 1), 4), dtype=t.float32, device=d.device) if want_Z else None
    h = (C.c_int64 * 2)()
    ws = L.workspace(lib.eac_dense_sl_workspace_bytes(n, code, 1 if condensed else 0))
    L.check(lib.eac_full_sl_lifetime(L.ptr(d), code, n, 1 if condensed else 0, int(n_clusters),
                                     L.ptr(Z), L.ptr(labels), h, L.ptr(ws), ws.numel(),
                                     L.stream_ptr()))
    nc = int(h[0])
    return nc, (None if h[1] else labels), (Z[: n - 1] if want_Z else None)


def full_sl_lifetime(mat, n_samples, n_clusters=0):
    """eac.py:337-364: (nc_stable, labels); empty int32 labels for one
    cluster. Raises ValueError where the reference's lifetime rule does."""
    nc, labels, _ = full_sl_lifetime_device(mat, n_samples, n_clusters)
    if labels is None:
        return nc, np.empty(0, dtype=np.int32)
    return nc, labels.cpu().numpy()


class EAC_FULL:
    """full.py:8-131 interface; the matrix is a device tensor."""

    def __init__(self, n_samples, dtype=np.uint8, condensed=False, **kwargs):
        self.n_samples = n_samples
        self.dtype = dtype
        self.nnz = 0
        self.n_updates = 0
        self.condensed = condensed
        self._validate_func_options()
        self.is_diassoc = False
        self.max_val = kwargs.get("max_val", None)
        t = L.torch()
        self._tdt = {np.dtype(np.uint8): t.uint8, np.dtype(np.uint16): t.uint16,
                     np.dtype(np.int32): t.int32, np.dtype(np.float32): t.float32}[
            np.dtype(self.dtype)]
        n = n_samples
        shape = (n * (n - 1) // 2,) if condensed else (n, n)
        self.coassoc = t.zeros(shape, dtype=self._tdt, device=L.device())

    def _validate_func_options(self):
        if type(self.n_samples) != int:
            raise TypeError("number of samples must be an integer")
        if self.n_samples <= 0:
            raise ValueError("number of samples must be an integer above 0")
        if not issubclass(self.dtype, (np.integer, np.floating)):
            raise TypeError("dtype must be a NumPy integer or float type")
        if np.dtype(self.dtype) not in _DT:
            raise TypeError("dtype must be uint8, uint16, int32 or float32 here")
        if self.condensed not in (True, False):
            raise ValueError("condensed must be boolean")

    def update_ensemble(self, ensemble):
        """All partitions (list-of-clusters format, partition.py:73-86) at
        once; counts add to the current matrix."""
        from .sparse import labels_from_ensemble
        labels, _ = labels_from_ensemble(ensemble, self.n_samples)
        self.update_labels(labels)

    def update_labels(self, labels):
        """Device entry: labels [P, n] (-1 = not listed)."""
        add, _ = coassoc_dense_device(labels, condensed=self.condensed, dtype=self.dtype)
        self.coassoc += add  # uint8 wraps, as the reference's in-place += does
        self.n_updates += int(np.prod(np.shape(labels)[:-1])) if np.ndim(labels) > 1 else 1

    def todense(self):
        """full.py:68-72: the n x n matrix (host)."""
        from scipy.spatial.distance import squareform
        m = self.coassoc.cpu().numpy()
        return squareform(m, checks=False) if self.condensed else m.copy()

    def tocsr(self):
        import scipy.sparse as sp
        return sp.csr_matrix(self.todense())

    def get_degree(self):
        """full.py:83-95: non-zero associations per row and nnz."""
        t = L.torch()
        lib = L.load_library()
        n = self.n_samples
        degree = t.empty(n, dtype=t.int32, device=self.coassoc.device)
        nnz = C.c_int64()
        ws = L.workspace(256)
        L.check(lib.eac_dense_degree(L.ptr(self.coassoc), _wcode(self.coassoc), n,
                                     1 if self.condensed else 0, L.ptr(degree), C.byref(nnz),
                                     L.ptr(ws), ws.numel(), L.stream_ptr()))
        self.degree = degree.cpu().numpy()
        self.nnz = nnz.value

    def make_diassoc(self, max_val=None):
        """full.py:98-112: coassoc <- max_val - coassoc (max_val = largest
        count + 1 unless given)."""
        if max_val is None:
            if self.max_val is not None:
                max_val = self.max_val + 1
            else:
                max_val = int(self.coassoc.max().item()) + 1
        lib = L.load_library()
        L.check(lib.eac_dense_diassoc(L.ptr(self.coassoc), self.coassoc.numel(),
                                      _wcode(self.coassoc), float(max_val), L.stream_ptr()))
        self.is_diassoc = True
