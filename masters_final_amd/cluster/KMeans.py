"""K-Means on the GPU behind the reference's API (MyML/MyML/cluster/KMeans.py).

K_Means keeps the reference's constructor and fit() contract (KMeans.py:26-201):
float32 data (anything else raises Warning, KMeans.py:119-122), 'random'
initial centroids drawn with random.sample from the data (KMeans.py:191-201)
or a given array, the label -> iteration count -> convergence check ->
centroid update loop with tol on |sum(dists) - previous| (KMeans.py:154-186).
Labeling and the centroid update run in libeac_hip.so; `precision` picks the
labeling arithmetic:
  "bf16x3" (default): split-BF16 MFMA distances, FP64 re-scoring of the points
           a rigorous bound cannot certify -> labels equal the FP64 oracle's;
  "fp64":  FP64 direct difference (the oracle's arithmetic);
  "fp32":  plain FP32 direct difference (the reference GPU kernel's spirit,
           KMeans.py:586-629).
The reference's label_mode / centroid_mode keywords are accepted; every mode
runs on the GPU.
"""
from __future__ import annotations

import ctypes as C
from random import sample

import numpy as np

from .. import _lib as L

EAC_KM_F32, EAC_KM_F64, EAC_KM_BF16X3, EAC_KM_F32_NUMPY = 0, 1, 2, 3
_PREC = {"fp32": EAC_KM_F32, "fp64": EAC_KM_F64, "bf16x3": EAC_KM_BF16X3,
         "numpy32": EAC_KM_F32_NUMPY}


def _koff(ks):
    off = np.zeros(len(ks) + 1, dtype=np.int64)
    np.cumsum(np.asarray(ks, dtype=np.int64), out=off[1:])
    return off


def label_device(X, centroids, k_list=None, precision="bf16x3"):
    """Label n points against P centroid sets. centroids: (sum k, d) with
    k_list giving each set's size (default: one set). Returns device
    (labels [P, n] int32, dists [P, n] float32, n_rescored)."""
    t = L.torch()
    lib = L.load_library()
    dX = L.as_device(X, t.float32)
    dC = L.as_device(centroids, t.float32)
    n, d = dX.shape
    ks = [dC.shape[0]] if k_list is None else list(k_list)
    P = len(ks)
    koff = _koff(ks)
    if koff[-1] != dC.shape[0]:
        raise ValueError("k_list does not match the number of centroids")
    labels = t.empty((P, n), dtype=t.int32, device=dX.device)
    dists = t.empty((P, n), dtype=t.float32, device=dX.device)
    h = (C.c_int64 * 2)()
    ws = L.workspace(lib.eac_kmeans_workspace_bytes(n, d, int(koff[-1]), P))
    L.check(lib.eac_kmeans_label(L.ptr(dX), n, d, L.ptr(dC), koff.ctypes.data_as(C.c_void_p), P,
                                 _PREC[precision], L.ptr(labels), L.ptr(dists), h, L.ptr(ws),
                                 ws.numel(), L.stream_ptr()))
    return labels, dists, int(h[0])


def update_device(X, labels, dists, k_list, d=None):
    """New centroids (sum k, d) from labels/dists [P, n] (KMeans.py:709-754)."""
    t = L.torch()
    lib = L.load_library()
    dX = L.as_device(X, t.float32)
    n, dd = dX.shape
    dl = L.as_device(labels, t.int32).reshape(-1, n)
    ddist = L.as_device(dists, t.float32).reshape(-1, n)
    ks = list(k_list)
    P = len(ks)
    koff = _koff(ks)
    out = t.empty((int(koff[-1]), dd), dtype=t.float32, device=dX.device)
    h = (C.c_int64 * 2)()
    ws = L.workspace(lib.eac_kmeans_workspace_bytes(n, dd, int(koff[-1]), P))
    L.check(lib.eac_kmeans_update(L.ptr(dX), n, dd, L.ptr(dl), L.ptr(ddist),
                                  koff.ctypes.data_as(C.c_void_p), P, L.ptr(out), h, L.ptr(ws),
                                  ws.numel(), L.stream_ptr()))
    return out, int(h[0])


def fit_batched(X, init_centroids, k_list, max_iter=3, tol=1e-4, precision="bf16x3"):
    """Batched K_Means.fit: P partitions of the same data in one call.
    Returns dict(labels [P, n], dists [P, n], centroids (sum k, d), iters [P],
    n_rescored, n_reseeded) with device tensors."""
    t = L.torch()
    lib = L.load_library()
    dX = L.as_device(X, t.float32)
    n, d = dX.shape
    cen = L.as_device(init_centroids, t.float32).clone()
    ks = list(k_list)
    P = len(ks)
    koff = _koff(ks)
    if koff[-1] != cen.shape[0]:
        raise ValueError("k_list does not match the number of centroids")
    labels = t.empty((P, n), dtype=t.int32, device=dX.device)
    dists = t.empty((P, n), dtype=t.float32, device=dX.device)
    iters = np.zeros(P, dtype=np.int32)
    h = (C.c_int64 * 2)()
    ws = L.workspace(lib.eac_kmeans_workspace_bytes(n, d, int(koff[-1]), P))
    L.check(lib.eac_kmeans_fit(L.ptr(dX), n, d, L.ptr(cen), koff.ctypes.data_as(C.c_void_p), P,
                               int(max_iter), float(tol), _PREC[precision], L.ptr(labels),
                               L.ptr(dists), iters.ctypes.data_as(C.c_void_p), h, L.ptr(ws),
                               ws.numel(), L.stream_ptr()))
    return {"labels": labels, "dists": dists, "centroids": cen, "iters": iters,
            "n_rescored": int(h[0]), "n_reseeded": int(h[1])}


class BatchedFitGraph:
    """fit_batched recorded once in a HIP graph (torch.cuda.CUDAGraph) and
    replayed: the batched ensemble of partition.py:117-122 (one K_Means.fit
    per partition, KMeans.py:117-189) with no host round trip. The data X,
    the cluster counts, max_iter, tol and the precision are fixed at
    capture; each run() copies new initial centroids into the graph's input
    buffer and replays. Results equal fit_batched's (exactly max_iter
    labelings run; a partition that stops keeps its centroids, so the
    extra labelings repeat its labels). Buffers are owned by the object
    (not the shared workspace cache) so that replays never see them move."""

    def __init__(self, X, k_list, max_iter=3, tol=1e-4, precision="bf16x3", capture=True):
        t = L.torch()
        lib = L.load_library()
        self.X = L.as_device(X, t.float32).contiguous()
        n, d = self.X.shape
        self.ks = list(k_list)
        P = len(self.ks)
        self.koff = _koff(self.ks)
        ktot = int(self.koff[-1])
        dev = self.X.device
        self.n, self.d, self.P, self.max_iter = n, d, P, int(max_iter)
        self.d_koff = t.from_numpy(self.koff).to(dev)
        # the warm-up call before capture starts from data rows (tied, all
        # equal centroids would flag every point for re-scoring)
        self.init = self.X.index_select(0, t.arange(ktot, device=dev) % n).contiguous()
        self.cen = t.zeros((ktot, d), dtype=t.float32, device=dev)
        self.labels = t.empty((P, n), dtype=t.int32, device=dev)
        self.dists = t.empty((P, n), dtype=t.float32, device=dev)
        self.iters = t.zeros(P, dtype=t.int32, device=dev)
        self.info = t.zeros(2, dtype=t.int64, device=dev)
        self.err = t.zeros(1, dtype=t.int32, device=dev)
        self.ws = t.empty(lib.eac_kmeans_fit_graph_workspace_bytes(n, d, ktot, P),
                          dtype=t.uint8, device=dev)
        self._args = (float(tol), _PREC[precision])
        side = t.cuda.Stream()
        side.wait_stream(t.cuda.current_stream())
        with t.cuda.stream(side):  # one eager call before capture
            self._enqueue(lib)
        t.cuda.current_stream().wait_stream(side)
        t.cuda.synchronize()
        self.graph = None
        if capture:  # (capture=False: a larger graph records _enqueue itself)
            self.graph = t.cuda.CUDAGraph()
            with t.cuda.graph(self.graph):
                self._enqueue(lib)

    def _enqueue(self, lib):
        tol, mode = self._args
        self.cen.copy_(self.init)
        L.check(lib.eac_kmeans_fit_graph(
            L.ptr(self.X), self.n, self.d, L.ptr(self.cen),
            self.koff.ctypes.data_as(C.c_void_p), L.ptr(self.d_koff), self.P, self.max_iter,
            tol, mode, L.ptr(self.labels), L.ptr(self.dists), L.ptr(self.iters),
            L.ptr(self.info), L.ptr(self.err), L.ptr(self.ws), self.ws.numel(), L.stream_ptr()))

    def run(self, init_centroids):
        """Replays the graph from these initial centroids (sum k, d); the
        results stay in the object's device buffers (labels, dists, cen,
        iters, info, err) until the next run."""
        t = L.torch()
        self.init.copy_(L.as_device(init_centroids, t.float32))
        self.graph.replay()
        return self

    def result(self):
        """dict as fit_batched returns (reads the error word: synchronises);
        raises when the replay hit a limit of the device-side path."""
        err = int(self.err.item())
        if err:
            raise RuntimeError(f"captured fit reported error bits {err} (flag list overflow = 1, "
                               f"more than 2048 empty clusters in a partition = 2)")
        info = self.info.cpu().numpy()
        return {"labels": self.labels, "dists": self.dists, "centroids": self.cen,
                "iters": self.iters.cpu().numpy(), "n_rescored": int(info[0]),
                "n_reseeded": int(info[1])}


def sum_f32_device(x):
    """numpy's float32 sum of each row of a 2-D device tensor (1-D: one row),
    bit-equal to ndarray.sum(): the inertia of KMeans.py:164."""
    t = L.torch()
    lib = L.load_library()
    dx = L.as_device(x, t.float32)
    x2 = dx.reshape(1, -1) if dx.dim() == 1 else dx
    rows, n = x2.shape
    out = t.empty(rows, dtype=t.float32, device=dx.device)
    ws = L.workspace(lib.eac_sum_f32_workspace_bytes(n, rows))
    L.check(lib.eac_sum_f32_rows(L.ptr(x2), n, rows, L.ptr(out), L.ptr(ws), ws.numel(),
                                 L.stream_ptr()))
    return out


def numba_label(data, centroids, dists, converge):
    """KMeans.py:453-490 (numba_label = njit(py_label)): FP64 direct difference,
    lowest k on ties; fills dists (float32) when converge. Runs on the GPU."""
    labels, d, _ = label_device(data, centroids, precision="fp64")
    if converge:
        dists[...] = d[0].cpu().numpy()
    return labels[0].cpu().numpy()


py_label = numba_label


def np_label(data, centroids, dists, converge):
    """KMeans.py:493-522: labels in numpy's float32 arithmetic
    (((data - c) ** 2).sum(axis=1) per centroid, strict <, lowest k on ties),
    computed on the GPU (EAC_KM_F32_NUMPY). As in the reference, `dists` is
    left untouched: its `dists = best_dist` (KMeans.py:519-520) rebinds a
    local name, so the caller never sees the distances."""
    labels, _, _ = label_device(data, centroids, precision="numpy32")
    return labels[0].cpu().numpy()


def numba_recompute_centroids(data, centroids, labels, dists):
    """KMeans.py:709-754 on the GPU."""
    k = np.asarray(centroids).shape[0] if not L.is_torch(centroids) else centroids.shape[0]
    out, _ = update_device(data, labels, dists, [k])
    return out.cpu().numpy()


py_recompute_centroids = numba_recompute_centroids


class K_Means:
    """KMeans.py:26-445 interface; compute in libeac_hip.so."""

    def __init__(self, n_clusters=8, tol=1e-4, max_iter=300, init="random", **kwargs):
        self.n_clusters = n_clusters
        self._label_mode = kwargs.get("label_mode", "cuda")
        self._centroid_mode = kwargs.get("centroid_mode", "numba")
        acceptable = {"label": ["python", "numpy", "numba", "cuda"],
                      "centroid": ["python", "numpy", "numba"]}
        if self._label_mode not in acceptable["label"]:
            raise ValueError("label_mode should be one of {}".format(acceptable["label"]))
        if self._centroid_mode not in acceptable["centroid"]:
            raise ValueError("centroid_mode should be one of {}".format(acceptable["centroid"]))
        self.precision = kwargs.get("precision", "bf16x3")
        if self.precision not in _PREC:
            raise ValueError(f"precision should be one of {sorted(_PREC)}")
        if isinstance(init, str) and init == "random":
            self.centroid_init_mode = "random"
        elif isinstance(init, np.ndarray):
            if init.shape[0] != n_clusters:
                raise Exception("Number of clusters indicated different from number of "
                                "centroids supplied.")
            self.centroid_init_mode = init.copy()
        else:
            raise ValueError("Centroid init may be 'random' or an ndarray containing the "
                             "centroids to use.")
        self.tol = tol
        self.max_iter = max_iter
        self._converge = kwargs.get("converge", True)
        self.inertia_ = np.inf
        self.iters_ = 0
        self.n_rescored_ = 0

    def _init_centroids(self, data):
        random_init = sample(range(self.N), self.n_clusters)  # KMeans.py:196
        return data[random_init]

    def fit(self, data):
        if L.is_torch(data):
            if data.dtype != L.torch().float32:
                raise Warning("DATA DUPLICATION: data converted to float32.")
            host_data = None
        else:
            if data.dtype != np.float32:
                raise Warning("DATA DUPLICATION: data converted to float32."
                              "TODO: accept other formats")
            host_data = data
        self.N, self.D = data.shape
        if isinstance(self.centroid_init_mode, str):
            src = host_data if host_data is not None else data.cpu().numpy()
            self.centroids = self._init_centroids(src)
        else:
            if self.centroid_init_mode.shape[0] != self.n_clusters:
                raise ValueError("Number of clusters indicated different from number of "
                                 "centroids supplied.")
            self.centroids = self.centroid_init_mode.copy()
        tol = self.tol if self._converge else -1.0
        r = fit_batched(data, self.centroids, [self.n_clusters], max_iter=self.max_iter, tol=tol,
                        precision=self.precision)
        self.labels_ = r["labels"][0].cpu().numpy()
        self._dists = r["dists"][0].cpu().numpy()
        self.cluster_centers_ = r["centroids"].cpu().numpy()
        self.centroids = self.cluster_centers_
        self.iters_ = int(r["iters"][0])
        # KMeans.py:164: the inertia is numpy's float32 sum of the last dists
        # (the device stopping rule used the same float32 sum, eac_sum_f32_rows)
        self.inertia_ = self._dists.sum() if self._converge else np.inf
        self.n_rescored_ = r["n_rescored"]
        return self
