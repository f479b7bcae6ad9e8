"""Evidence-accumulation clustering driver pieces on the GPU, behind the
reference's API (MyML/MyML/EAC/eac.py).

sp_sl_lifetime replaces eac.py:199-267 (Kruskal minimum spanning tree,
lifetime rule of eac.py:366-391, cut, scipy connected components). The
forest is built by Boruvka on the GPU (eac_sl_lifetime, rule EAC_CUT_EAC);
the partition equals the reference's: with the lifetime rule the cut removes
exactly the forest edges heavier than the lower end of the largest gap,
which does not depend on which of several equal-weight spanning trees is
taken. Labels are numbered by first occurrence, as scipy's
connected_components numbers them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib as L
from ..graph.singlelink import EAC_CUT_EAC, lifetime_cut


def _csr_parts(mat):
    """(indptr, indices, data) of a scipy CSR matrix or a tuple of arrays/tensors."""
    if isinstance(mat, tuple):
        return mat
    return mat.indptr, mat.indices, mat.data


def sp_sl_lifetime_device(mat, max_dist=None, n_clusters=0):
    """Device variant: (n_comps, labels tensor or None for one cluster, info)."""
    indptr, indices, data = _csr_parts(mat)
    labels, info = lifetime_cut(indices, data, indptr, None, rule=EAC_CUT_EAC,
                                max_dist=max_dist, n_clusters=n_clusters, symmetric=False)
    return info["n_components"], labels, info


def sp_sl_lifetime(mat, max_dist=None, n_clusters=0):
    """eac.py:199-267: (n_fclusts, labels) for a CSR co-association/dissimilarity
    graph (any orientation: every stored entry is an undirected edge). With one
    final cluster the labels are an empty int32 array, as in eac.py:263-265."""
    n_comps, labels, _ = sp_sl_lifetime_device(mat, max_dist, n_clusters)
    if labels is None:
        return 1, np.empty(0, dtype=np.int32)
    return n_comps, labels.cpu().numpy()


class EAC:
    """eac.py:40-179 interface. With sparse=True, buildMatrix counts the
    ensemble's co-associations on the GPU (EAC_CSR: exact, or the
    reference's capacity-capped matrix with sparse_exact=False) and turns
    them into dissimilarities (max + 1) - count; finalClustering runs the
    Boruvka forest, the lifetime rule and the cut on the GPU
    (sp_sl_lifetime). With sparse=False (the reference's default) the dense
    EAC_FULL matrix and full_sl_lifetime run instead (eac.py:139-144,
    :166-169), on the GPU as well. The disk-backed Kruskal (eac.py:270-335)
    is not built."""

    def __init__(self, n_samples, **kwargs):
        self.n_samples = n_samples
        self.n_partitions = kwargs.get("n_partitions", 100)
        self.iters = kwargs.get("iters", 3)
        self.n_clusters = kwargs.get("n_clusters", "sqrt")
        self.toFiles = False
        self.toFiles_folder = None
        self.condensed = kwargs.get("condensed", True)
        self.kNN = kwargs.get("kNN", False)
        self.assoc_dtype = kwargs.get("assoc_dtype", np.uint8)
        self.sparse = kwargs.get("sparse", False)
        self.sp_sort = kwargs.get("sparse_sort_mode", "surgical")
        self.sp_max_assocs = kwargs.get("sparse_max_assocs", None)
        self.sp_max_assocs_factor = kwargs.get("sparse_max_assocs_factor", 3)
        self.sp_max_assocs_mode = kwargs.get("sparse_max_assocs_mode", "linear")
        self.sp_keep_degree = kwargs.get("sparse_keep_degree", False)
        # False: the reference's capacity-capped (lossy) matrix, for parity
        # studies; True (default): exact counts
        self.sp_exact = kwargs.get("sparse_exact", True)
        self.full = not self.sparse and not self.kNN
        self.coassoc_store = kwargs.get("coassoc_store", False)
        self.linkage = kwargs.get("linkage", "SL")
        self.disk_mst = kwargs.get("sl_disk", False)

    def _validate_params(self):
        pass

    def generateEnsemble(self):
        pass

    def buildMatrix(self, ensemble):
        from .sparse import EAC_CSR
        if self.kNN:
            raise NotImplementedError("kNN matrix building has not been included in this "
                                      "version.")
        if self.full:
            # eac.py:139-144: dense counts, degrees, dissimilarities
            from .full import EAC_FULL
            coassoc = EAC_FULL(self.n_samples, condensed=self.condensed,
                               dtype=self.assoc_dtype)
            coassoc.update_ensemble(ensemble)
            coassoc.get_degree()
            coassoc.make_diassoc()
            self.coassoc = coassoc
            return
        if self.sp_max_assocs is None:
            self.sp_max_assocs = max(max(map(np.size, p)) for p in ensemble)
            self.sp_max_assocs *= self.sp_max_assocs_factor
        mode = self.sp_max_assocs_mode if self.condensed else "constant"
        coassoc = EAC_CSR(self.n_samples, max_assocs=int(self.sp_max_assocs),
                          condensed=self.condensed, max_assocs_type=mode,
                          sort_mode=self.sp_sort, dtype=self.assoc_dtype,
                          exact=self.sp_exact)
        coassoc.update_ensemble(ensemble)
        coassoc._condense(keep_degree=self.sp_keep_degree)
        coassoc.make_diassoc()
        self.coassoc = coassoc

    def buildMatrixFromLabels(self, labels, k_list):
        """Device entry: labels [P, n] of the ensemble (no list conversion)."""
        from .sparse import EAC_CSR
        coassoc = EAC_CSR(self.n_samples, max_assocs=1, condensed=self.condensed,
                          max_assocs_type="constant", dtype=self.assoc_dtype)
        coassoc.update_labels(labels, k_list)
        coassoc._condense()
        coassoc.make_diassoc()
        self.coassoc = coassoc

    def finalClustering(self, n_clusters=0):
        c = self.coassoc
        if self.full:
            # eac.py:166-169: full_sl_lifetime on the dense dissimilarities
            from .full import full_sl_lifetime_device
            n_fclusts, labels, _ = full_sl_lifetime_device(c.coassoc, self.n_samples,
                                                           n_clusters=n_clusters)
        else:
            n_fclusts, labels, _ = sp_sl_lifetime_device((c.indptr, c.indices, c.data),
                                                         self.n_partitions,
                                                         n_clusters=n_clusters)
        self.n_fclusts = n_fclusts
        self.labels = (np.empty(0, dtype=np.int32) if labels is None
                       else labels.cpu().numpy())
        return self.labels


class EnsembleCoassocGraph:
    """The EAC hot path up to the dissimilarity matrix recorded in one HIP
    graph: the batched K-Means ensemble (eac_kmeans_fit_graph) and the
    one-pass co-association with (max + 1) - count dissimilarities
    (eac_coassoc_graph, condensed rows), with no host round trip inside.
    X, the cluster counts, iters and the precision are fixed at capture;
    run(init) replays from new initial centroids and reads the matrix size
    back once. cap_nnz (default: every condensed pair, n (n - 1) / 2)
    bounds the matrix; the staging pool holds cap_nnz entries as well."""

    def __init__(self, X, k_list, iters=3, tol=1e-4, precision="bf16x3", cap_nnz=None,
                 pad_tail=False):
        from ..cluster.KMeans import BatchedFitGraph
        t = L.torch()
        lib = L.load_library()
        self.fit = BatchedFitGraph(X, k_list, max_iter=iters, tol=tol, precision=precision,
                                   capture=False)
        f = self.fit
        n, P = f.n, f.P
        cap = int(cap_nnz) if cap_nnz else max(n * (n - 1) // 2, 1)
        dev = f.X.device
        self.cap = cap
        # pad_tail: the matrix as a graph of n + 1 vertices whose last one
        # owns the unused capacity as self loops (indptr[n + 1] = cap, the
        # entries past nnz set on the device after each build), the input
        # shape of a recorded lifetime cut. The loops carry the largest
        # dissimilarity (the max count): the cut's weight-level split samples
        # weights evenly over all entries, and loops at the smallest weight
        # would pull its light level below every real edge.
        self.pad_tail = bool(pad_tail)
        self.indptr_ext = t.empty(n + 2, dtype=t.int64, device=dev)
        self.indptr_ext[n + 1] = cap
        self.indptr = self.indptr_ext[:n + 1]
        self.indices = t.empty(cap, dtype=t.int32, device=dev)
        self.data = t.empty(cap, dtype=t.int32, device=dev)
        self.out = t.zeros(4, dtype=t.int64, device=dev)
        nb = lib.eac_coassoc_stage_bytes(n, 0, P, cap)
        self.stage = t.empty(nb, dtype=t.uint8, device=dev)
        self.cws = t.empty(lib.eac_coassoc_workspace_bytes(n, P, int(f.koff[-1])), dtype=t.uint8,
                           device=dev)
        side = t.cuda.Stream()
        side.wait_stream(t.cuda.current_stream())
        with t.cuda.stream(side):
            self._enqueue_coassoc(lib)
        t.cuda.current_stream().wait_stream(side)
        t.cuda.synchronize()
        if self.pad_tail:
            self._slot = t.arange(cap, dtype=t.int64, device=dev)
        self.graph = t.cuda.CUDAGraph()
        with t.cuda.graph(self.graph):
            f._enqueue(lib)
            self._enqueue_coassoc(lib)
            if self.pad_tail:
                past = self._slot >= self.indptr[n]
                self.indices.masked_fill_(past, n)
                top = self.out[1].to(t.int32)  # on the device: no read-back
                self.data.copy_(t.where(past, top, self.data))

    def _enqueue_coassoc(self, lib):
        f = self.fit
        L.check(lib.eac_coassoc_graph(
            L.ptr(f.labels), f.n, f.koff.ctypes.data_as(C.c_void_p), L.ptr(f.d_koff), f.P, 0, -1,
            L.ptr(self.indptr), L.ptr(self.indices), L.ptr(self.data), self.cap,
            L.ptr(self.stage), self.stage.numel(), L.ptr(self.out), L.ptr(self.cws),
            self.cws.numel(), L.stream_ptr()))

    def run(self, init_centroids):
        """Replays from these initial centroids; returns (indptr, indices,
        data, max_count, fit result) as eac_pipeline's matrix."""
        t = L.torch()
        self.fit.init.copy_(L.as_device(init_centroids, t.float32))
        self.graph.replay()
        nnz, mx, err, _ = (int(v) for v in self.out.cpu())
        if err:
            raise RuntimeError(f"captured co-association reported error bits {err} (pool too "
                               f"small = 1, more than {self.cap} entries = 2, label >= k = 4)")
        fit = self.fit.result()
        return self.indptr, self.indices[:nnz], self.data[:nnz], mx, fit


class EACPipelineGraph:
    """The whole EAC hot path with no host round trip before its one
    read-back: an EnsembleCoassocGraph replay (ensemble, co-association,
    dissimilarities) followed on the same stream by the lifetime cut
    recorded as a batched-cut HIP graph (eac_sl_lifetime_batched_graph_*:
    the matrix as graph 0, a padding vertex that owns the unused matrix
    capacity as self loops as graph 1, whose entries past nnz are set on the
    device inside the first graph). The capacity is the nnz of one eager
    pass from `init` times `margin`; a replay whose matrix outgrows it is
    reported (error bit 2), never cut. The cut's round count is tuned on the
    first replay (its rounds plus `round_margin`; a replay that needs more
    is detected on the device and run again at the bound)."""

    def __init__(self, X, k_list, init, iters=3, tol=1e-4, precision="bf16x3", n_clusters=0,
                 margin=1.25, round_margin=3):
        t = L.torch()
        self._lib = L.load_library()
        first = eac_pipeline(X, k_list, iters=iters, precision=precision, init=init,
                             n_clusters=n_clusters)
        cap = int(first["nnz"] * margin) + 4096
        self.coassoc = EnsembleCoassocGraph(X, k_list, iters=iters, tol=tol, precision=precision,
                                            cap_nnz=cap, pad_tail=True)
        g = self.coassoc
        n, P = g.fit.n, g.fit.P
        dev = g.fit.X.device
        self.n, self.P, self.n_clusters = n, P, int(n_clusters)
        self.goff = t.tensor([0, n, n + 1], dtype=t.int64, device=dev)
        self.labels = t.zeros(n + 1, dtype=t.int32, device=dev)
        self.gout = t.zeros((2, 4), dtype=t.int64, device=dev)
        self.summary = t.zeros(8, dtype=t.int64, device=dev)
        nb = self._lib.eac_sl_lifetime_batched_workspace_bytes(n + 1, cap, 2)
        self.ws = t.empty(int(nb), dtype=t.uint8, device=dev)
        self._h = None
        self.bound_rounds = int(self._lib.eac_mst_round_bound(n + 1, cap, 0))
        self._capture(0)
        self.run(init)  # the first replay tunes the round count
        want = min(self._last_rounds + int(round_margin), self.bound_rounds)
        if want < self.rounds:
            self._capture(want)

    def _capture(self, max_rounds):
        from ..graph.singlelink import EAC_CUT_EAC
        self.close()
        g = self.coassoc
        h = C.c_void_p()
        L.torch().cuda.current_stream().synchronize()
        L.check(self._lib.eac_sl_lifetime_batched_graph_create(
            L.ptr(g.indices), L.ptr(g.data), L.weight_code(g.data.dtype), L.ptr(g.indptr_ext),
            None, self.n + 1, g.cap, 0, L.ptr(self.goff), 2, EAC_CUT_EAC, float(self.P),
            self.n_clusters, int(max_rounds), L.ptr(self.labels), L.ptr(self.gout),
            L.ptr(self.summary), L.ptr(self.ws), self.ws.numel(), C.byref(h)))
        self._h = h
        self.rounds = int(self._lib.eac_sl_lifetime_batched_graph_rounds(h))

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.eac_sl_lifetime_batched_graph_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def run(self, init_centroids):
        """Replays the pipeline from these initial centroids; returns
        eac_pipeline's dict (its labels a copy; its ensemble labels the
        recorded graph's buffer, overwritten by the next replay)."""
        t = L.torch()
        g = self.coassoc
        g.fit.init.copy_(L.as_device(init_centroids, t.float32))
        g.graph.replay()
        L.check(self._lib.eac_sl_lifetime_batched_graph_launch(self._h, L.stream_ptr()))
        head = t.cat([g.out, self.summary[:5], self.gout[0]]).cpu().tolist()
        nnz, mx, err = int(head[0]), int(head[1]), int(head[2])
        if err:
            raise RuntimeError(f"recorded co-association reported error bits {err} (pool too "
                               f"small = 1, more than {g.cap} entries = 2, label >= k = 4)")
        flags = int(head[8])
        if flags & 16 and self.rounds < self.bound_rounds:
            self._capture(0)  # the cut needed more rounds than recorded
            L.check(self._lib.eac_sl_lifetime_batched_graph_launch(self._h, L.stream_ptr()))
            head = t.cat([g.out, self.summary[:5], self.gout[0]]).cpu().tolist()
            flags = int(head[8])
        if flags:
            raise RuntimeError(f"recorded lifetime cut: device error flags {flags}")
        self._last_rounds = int(head[6])
        if head[12] < 0:
            raise ValueError("lifetime cut: the requested number of clusters cannot be cut "
                             "from the forest")
        nc, empty = int(head[9]), int(head[10])
        fit = g.fit.result()
        return {"n_clusters": nc, "labels": None if empty else self.labels[:self.n].clone(),
                "n_rescored": fit["n_rescored"], "nnz": nnz, "ensemble_labels": fit["labels"],
                "max_count": mx}


def eac_pipeline_graph(graph, init, n_clusters=0):
    """eac_pipeline with the ensemble and the co-association replayed from
    an EnsembleCoassocGraph, then the forest and the cut."""
    indptr, indices, data, mx, fit = graph.run(init)
    nc, lab, info = sp_sl_lifetime_device((indptr, indices, data), graph.fit.P,
                                          n_clusters=n_clusters)
    return {"n_clusters": nc, "labels": lab, "n_rescored": fit["n_rescored"],
            "nnz": int(indices.numel()), "ensemble_labels": fit["labels"], "max_count": mx}


def eac_pipeline(X, k_list, iters=3, seed=0, precision="bf16x3", init=None, n_clusters=0):
    """Whole EAC hot path on the GPU: batched K-Means ensemble -> exact
    co-association -> dissimilarity (max + 1) - count -> Boruvka forest ->
    lifetime cut (max_dist = number of partitions, eac.py:159-161) -> labels.
    Returns dict(n_clusters, labels (device tensor or None for one cluster),
    n_rescored, nnz)."""
    from ..utils.partition import generate_ensemble_batched
    from .sparse import coassoc_device
    labels, fit = generate_ensemble_batched(X, k_list, iters=iters, seed=seed,
                                            precision=precision, init=init)
    indptr, indices, data, mx = coassoc_device(labels, k_list, full_rows=False,
                                               w_offset="max+1")
    nc, lab, info = sp_sl_lifetime_device((indptr, indices, data), len(k_list),
                                          n_clusters=n_clusters)
    return {"n_clusters": nc, "labels": lab, "n_rescored": fit["n_rescored"],
            "nnz": int(indices.numel()), "ensemble_labels": labels, "max_count": mx}
