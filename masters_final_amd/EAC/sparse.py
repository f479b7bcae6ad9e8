"""Sparse co-association matrix on the GPU behind the reference's API
(MyML/MyML/EAC/sparse.py).

EAC_CSR keeps the reference's constructor, update_ensemble / _condense /
make_diassoc / tocsr / todense sequence (sparse.py:12-304). The counting runs
in libeac_hip.so (eac_coassoc_count + eac_coassoc_fill). By default it is
exact: the reference's per-row max_assocs capacity (indptr_constant /
indptr_linear) and its silent discards (nnz_discarded, sparse.py:501-503) do
not apply, so nnz_discarded is 0 and the counts equal the uncapped
reference. With exact=False the capacities apply as in the reference
(capped-compat mode: the same kept columns, counts and nnz_discarded as the
reference's surgical build, for parity studies against its lossy matrix).
Counts are int32 (the reference's uint8 default wraps at 255,
sparse.py:300). Arrays stay on the device (torch tensors); .csr builds a
scipy matrix on request.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib as L


def labels_from_ensemble(ensemble, n_samples):
    """List-of-clusters partitions (partition.py:73-86 format) -> labels [P, n]
    int32 (-1 where a partition leaves a sample out) and k per partition."""
    P = len(ensemble)
    labels = np.full((P, n_samples), -1, dtype=np.int32)
    ks = []
    for p, part in enumerate(ensemble):
        for c, members in enumerate(part):
            labels[p, np.asarray(members, dtype=np.int64)] = c
        ks.append(max(len(part), 1))
    return labels, ks


ONE_PASS_MIN_N = 65536  # smaller ensembles: the count pass is cheap, two passes


def coassoc_device(labels, k_list, full_rows=False, w_offset=0, one_pass=None,
                   sample_stride=None, margin=1 / 8, info=None):
    """Co-association CSR of labels [P, n] on the GPU.
    Returns (indptr int64 [n+1], indices int32 [nnz], data int32 [nnz], max_count).
    data = count, or w_offset - count when w_offset != 0.
    one_pass (default: n >= ONE_PASS_MIN_N) visits the pairs once: a count
    over every sample_stride-th row (default n // 2048, at most 32) estimates
    nnz, the staged pass writes into a pool of estimate x (1 + margin)
    entries and a compaction builds the CSR; a pool that proves too small
    falls back to the two-pass build. info (a dict) receives the path taken."""
    t = L.torch()
    lib = L.load_library()
    dl = L.as_device(labels, t.int32)
    if dl.dim() == 1:
        dl = dl[None]
    P, n = dl.shape
    koff = np.zeros(P + 1, dtype=np.int64)
    np.cumsum(np.asarray(k_list, dtype=np.int64), out=koff[1:])
    indptr = t.empty(n + 1, dtype=t.int64, device=dl.device)
    nnz, mx = C.c_int64(), C.c_int32()
    ws = L.workspace(lib.eac_coassoc_workspace_bytes(n, P, int(koff[-1])))
    fr = 1 if full_rows else 0
    kp = koff.ctypes.data_as(C.c_void_p)
    if one_pass is None:
        one_pass = n >= ONE_PASS_MIN_N
    if info is not None:
        info["path"] = "two-pass"
    if one_pass:
        stride = int(sample_stride) if sample_stride else max(1, min(32, n // 2048))
        est = C.c_int64()
        L.check(lib.eac_coassoc_estimate(L.ptr(dl), n, kp, P, fr, stride, C.byref(est), L.ptr(ws),
                                         ws.numel(), L.stream_ptr()))
        pool = max(int(est.value * (1 + margin)), 0)
        nb = lib.eac_coassoc_stage_bytes(n, fr, P, pool)
        stage = t.empty(nb, dtype=t.uint8, device=dl.device)
        L.check(lib.eac_coassoc_stage(L.ptr(dl), n, kp, P, fr, L.ptr(indptr), L.ptr(stage), nb,
                                      C.byref(nnz), C.byref(mx), L.ptr(ws), ws.numel(),
                                      L.stream_ptr()))
        if info is not None:
            info.update(estimate=est.value, pool_entries=pool, sample_stride=stride)
        if nnz.value >= 0:
            m = max(nnz.value, 1)
            indices = t.empty(m, dtype=t.int32, device=dl.device)
            data = t.empty(m, dtype=t.int32, device=dl.device)
            if w_offset == "max+1":
                w_offset = mx.value + 1
            L.check(lib.eac_coassoc_compact(n, L.ptr(indptr), L.ptr(stage), nb, int(w_offset),
                                            L.ptr(indices), L.ptr(data), L.stream_ptr()))
            if info is not None:
                info["path"] = "one-pass"
            return indptr, indices[: nnz.value], data[: nnz.value], mx.value
        del stage
        if info is not None:
            info["path"] = "one-pass overflow, two-pass"
    L.check(lib.eac_coassoc_count(L.ptr(dl), n, koff.ctypes.data_as(C.c_void_p), P,
                                  1 if full_rows else 0, L.ptr(indptr), C.byref(nnz), C.byref(mx),
                                  L.ptr(ws), ws.numel(), L.stream_ptr()))
    indices = t.empty(max(nnz.value, 1), dtype=t.int32, device=dl.device)
    data = t.empty(max(nnz.value, 1), dtype=t.int32, device=dl.device)
    if w_offset == "max+1":
        w_offset = mx.value + 1
    L.check(lib.eac_coassoc_fill(L.ptr(dl), n, koff.ctypes.data_as(C.c_void_p), P,
                                 1 if full_rows else 0, L.ptr(indptr), int(w_offset),
                                 L.ptr(indices), L.ptr(data), L.ptr(ws), ws.numel(),
                                 L.stream_ptr()))
    return indptr, indices[: nnz.value], data[: nnz.value], mx.value


def coassoc_capped_device(labels, k_list, cap_off, full_rows=False, max_assocs=0):
    """The reference's capacity-capped co-association (EAC_CSR with
    max_assocs, sort_mode "surgical") on the GPU: the exact build with
    per-entry partition ranks (eac_coassoc_fill_ranked), then the per-row
    selection of the columns the reference keeps (eac_coassoc_cap).
    cap_off: int64 [n+1] capacity offsets (indptr_constant / indptr_linear).
    Returns (indptr, indices, data, max_count, nnz_discarded) with device
    tensors; raises ValueError where the reference's first partition writes
    past a row's capacity."""
    t = L.torch()
    lib = L.load_library()
    dl = L.as_device(labels, t.int32)
    if dl.dim() == 1:
        dl = dl[None]
    P, n = dl.shape
    koff = np.zeros(P + 1, dtype=np.int64)
    np.cumsum(np.asarray(k_list, dtype=np.int64), out=koff[1:])
    indptr = t.empty(n + 1, dtype=t.int64, device=dl.device)
    nnz, mx = C.c_int64(), C.c_int32()
    ws = L.workspace(lib.eac_coassoc_workspace_bytes(n, P, int(koff[-1])))
    fr = 1 if full_rows else 0
    L.check(lib.eac_coassoc_count(L.ptr(dl), n, koff.ctypes.data_as(C.c_void_p), P, fr,
                                  L.ptr(indptr), C.byref(nnz), C.byref(mx), L.ptr(ws),
                                  ws.numel(), L.stream_ptr()))
    m = max(nnz.value, 1)
    ind = t.empty(m, dtype=t.int32, device=dl.device)
    cnt = t.empty(m, dtype=t.int32, device=dl.device)
    rank = t.empty(m, dtype=t.int32, device=dl.device)
    L.check(lib.eac_coassoc_fill_ranked(L.ptr(dl), n, koff.ctypes.data_as(C.c_void_p), P, fr,
                                        L.ptr(indptr), L.ptr(ind), L.ptr(cnt), L.ptr(rank),
                                        L.ptr(ws), ws.numel(), L.stream_ptr()))
    d_off = None if full_rows else L.as_device(np.asarray(cap_off, dtype=np.int64), t.int64)
    cap_total = nnz.value if full_rows else min(nnz.value, int(np.asarray(cap_off)[-1]))
    out_ip = t.empty(n + 1, dtype=t.int64, device=dl.device)
    out_ind = t.empty(max(cap_total, 1), dtype=t.int32, device=dl.device)
    out_dat = t.empty(max(cap_total, 1), dtype=t.int32, device=dl.device)
    h = (C.c_int64 * 4)()
    ws2 = t.empty(lib.eac_coassoc_cap_workspace_bytes(n), dtype=t.uint8, device=dl.device)
    L.check(lib.eac_coassoc_cap(L.ptr(indptr), L.ptr(ind), L.ptr(cnt), L.ptr(rank), n, P, fr,
                                L.ptr(d_off), int(max_assocs), max(cap_total, 1),
                                L.ptr(out_ip), L.ptr(out_ind), L.ptr(out_dat), h, L.ptr(ws2),
                                ws2.numel(), L.stream_ptr()))
    kept = int(h[0])
    return out_ip, out_ind[:kept], out_dat[:kept], int(h[2]), int(h[1])


def indptr_constant(n_samples, max_assocs):
    """sparse.py:368-369 capacity profile: every row holds max_assocs."""
    return np.arange(n_samples + 1, dtype=np.int64) * max_assocs


def indptr_linear(n_samples, max_assocs, n_s, n_e, val_s, val_e):
    """sparse.py:371-413 capacity profile: val_s * max_assocs up to row
    n_s * n, then a line down to val_e * max_assocs at row n_e * n (float
    line values truncated to integers), returned as exclusive offsets."""
    x = np.arange(n_samples, dtype=np.int32)
    y = np.empty(n_samples + 1, dtype=np.int64)
    n_s, n_e = n_s * n_samples, n_e * n_samples
    val_s, val_e = val_s * max_assocs, val_e * max_assocs
    m = (val_e - val_s) / (n_e - n_s)
    b = val_s - n_s * m
    n_s, n_e, val_s, val_e = int(n_s), int(n_e), int(val_s), int(val_e)
    y[:n_s] = val_s
    y[n_s:n_e] = m * x[n_s:n_e] + b
    y[n_e:] = val_e
    out = np.zeros_like(y)
    out[1:] = np.cumsum(y[:-1])
    out[0] = 0
    return out


class EAC_CSR:
    """sparse.py:12-304 interface; GPU build, exact unless exact=False."""

    def __init__(self, n_samples, max_assocs, max_assocs_type="constant", dtype=np.uint8,
                 condensed=False, sort_mode="numpy", **kwargs):
        self.n_samples = n_samples
        self.max_assocs = max_assocs
        self.max_assocs_type = max_assocs_type
        self.dtype = dtype
        self.condensed = condensed
        self.sort_mode = sort_mode
        self._validate_func_options()
        # linear capacity profile arguments (sparse.py:57-61)
        self.n_s = kwargs.get("n_s", 0.05)
        self.n_e = kwargs.get("n_e", 1.0)
        self.val_s = kwargs.get("val_s", 1.0)
        self.val_e = kwargs.get("val_e", 0.05)
        self.exact = kwargs.get("exact", True)
        self.nnz = 0
        self.nnz_discarded = 0
        self.csr_form = False
        self.is_diassoc = False
        self.max_val = kwargs.get("max_val", None)
        self._labels = None
        self._ks = None
        self.indptr = self.indices = self.data = None
        self.max_count = 0

    def _validate_func_options(self):
        if self.condensed not in (True, False):
            raise TypeError("condensed must be boolean")
        if type(self.max_assocs) != int:
            raise TypeError("max_assocs must be an integer")
        if self.max_assocs <= 0:
            raise ValueError("max_assocs must be an integer above 0")
        if type(self.n_samples) != int:
            raise TypeError("number of samples must be an integer")
        if self.n_samples <= 0:
            raise ValueError("number of samples must be an integer above 0")
        if not issubclass(self.dtype, (np.integer, np.floating)):
            raise TypeError("dtype must be a NumPy integer or float type")
        if self.sort_mode not in ("numpy", "normal", "simple", "surgical"):
            raise ValueError("sort_mode must be one of the following: "
                             "('numpy', 'normal', 'simple', 'surgical')")
        if self.max_assocs_type not in ("constant", "linear"):
            raise ValueError("max_assocs_type must be one of the following: "
                             "('constant', 'linear')")
        if self.max_assocs_type != "constant" and not self.condensed:
            raise ValueError('max_assocs_type must be "constant" if matrix is condensed')

    def update_ensemble(self, ensemble, progress=False):
        """Count all partitions (list-of-clusters format) at once on the GPU."""
        labels, ks = labels_from_ensemble(ensemble, self.n_samples)
        self.update_labels(labels, ks)

    def capacity_offsets(self):
        """The reference's row capacities as offsets (_allocate_arrays,
        sparse.py:67-80)."""
        if self.max_assocs_type == "constant":
            return indptr_constant(self.n_samples, self.max_assocs)
        return indptr_linear(self.n_samples, self.max_assocs, self.n_s, self.n_e, self.val_s,
                             self.val_e)

    def update_labels(self, labels, k_list):
        """Device entry: labels [P, n] (tensor or array), k per partition."""
        self._labels = labels
        self._ks = list(k_list)
        if self.exact:
            self.indptr, self.indices, self.data, self.max_count = coassoc_device(
                labels, self._ks, full_rows=not self.condensed)
            self.nnz_discarded = 0
        else:
            (self.indptr, self.indices, self.data, self.max_count,
             self.nnz_discarded) = coassoc_capped_device(
                labels, self._ks, self.capacity_offsets(), full_rows=not self.condensed,
                max_assocs=self.max_assocs)
        self.nnz = int(self.indices.numel())
        self.csr_form = False

    def _condense(self, keep_degree=False):
        """The GPU build already produces packed rows (condense_eac_csr)."""
        self.csr_form = True
        self._csr = None

    def make_diassoc(self, max_val=None):
        """sparse.py:295-304: data <- (max + 1) - data, on the device."""
        if max_val is None:
            max_val = (self.max_val + 1) if self.max_val is not None else self.max_count + 1
        self.data = int(max_val) - self.data
        self.is_diassoc = True
        self._csr = None

    @property
    def csr(self):
        import scipy.sparse as sp
        if getattr(self, "_csr", None) is None:
            n = self.n_samples
            self._csr = sp.csr_matrix((self.data.cpu().numpy(), self.indices.cpu().numpy(),
                                       self.indptr.cpu().numpy()), shape=(n, n))
        return self._csr

    def tocsr(self, copy=False):
        return self.csr.copy() if copy else self.csr

    def todense(self):
        return self.csr.todense()
