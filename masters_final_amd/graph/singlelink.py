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


class LifetimeCutGraph:
    """lifetime_cut_batched recorded once as a hipGraph and replayed per batch.

    Capacities (n_cap vertices, m_cap CSR entries, g_cap graphs) fix the
    recorded shape. Each batch -- the disjoint union of lifetime_cut_batched
    -- is copied into persistent device buffers and padded: vertices past N
    form one extra graph of isolated vertices, whose last vertex owns the
    padding entries as self loops (self loops never join a forest), and the
    graphs after it are empty. So a batch needs N < n_cap, or N == n_cap and
    M == m_cap; and G < g_cap whenever N < n_cap.

    Replays run a fixed number of Boruvka rounds. max_rounds=0 records the
    bound that always suffices (L * (bits(n_cap - 1) + 3), L = 1 to 4 weight
    levels of the Borůvka split, by edges per vertex); with tune=True
    the first replay's round count plus `margin` is recorded instead. A batch
    that needs more rounds than recorded is detected on the device (summary
    flag 16) and run again on a graph recorded at the bound, so results never
    depend on the tuning.
    """

    def __init__(self, n_cap, m_cap, g_cap, weight_dtype, rule=EAC_CUT_EAC, max_dist=None,
                 n_clusters=0, symmetric=True, max_rounds=0, tune=False, margin=3):
        t = L.torch()
        self._lib = L.load_library()
        dev = L.device()
        if n_cap < 2 or m_cap < 1 or g_cap < 2:
            raise ValueError("LifetimeCutGraph: need n_cap >= 2, m_cap >= 1, g_cap >= 2")
        self.n_cap, self.m_cap, self.g_cap = int(n_cap), int(m_cap), int(g_cap)
        self.rule, self.symmetric = int(rule), bool(symmetric)
        self.max_dist = -1.0 if max_dist is None else float(max_dist)
        self.n_clusters = int(n_clusters)
        self.tune, self.margin = bool(tune), int(margin)
        self.dest = t.zeros(self.m_cap, dtype=t.int32, device=dev)
        self.weight = t.zeros(self.m_cap, dtype=weight_dtype, device=dev)
        self.indptr = t.zeros(self.n_cap + 1, dtype=t.int64, device=dev)
        self.goff = t.zeros(self.g_cap + 1, dtype=t.int64, device=dev)
        self.labels = t.zeros(self.n_cap, dtype=t.int32, device=dev)
        self.gout = t.zeros((self.g_cap, 4), dtype=t.int64, device=dev)
        self.summary = t.zeros(8, dtype=t.int64, device=dev)
        nb = self._lib.eac_sl_lifetime_batched_workspace_bytes(self.n_cap, self.m_cap, self.g_cap)
        self.ws = t.empty(int(nb), dtype=t.uint8, device=dev)
        self._h = None
        self._capture(max_rounds)
        self.bound_rounds = self._bound()

    def _bound(self):
        return int(self._lib.eac_mst_round_bound(self.n_cap, self.m_cap,
                                                  1 if self.symmetric else 0))

    def _capture(self, max_rounds):
        self.close()
        h = C.c_void_p()
        L.torch().cuda.current_stream().synchronize()
        L.check(self._lib.eac_sl_lifetime_batched_graph_create(
            L.ptr(self.dest), L.ptr(self.weight), L.weight_code(self.weight.dtype),
            L.ptr(self.indptr), None, self.n_cap, self.m_cap, 1 if self.symmetric else 0,
            L.ptr(self.goff), self.g_cap, self.rule, self.max_dist, self.n_clusters,
            int(max_rounds), L.ptr(self.labels), L.ptr(self.gout), L.ptr(self.summary),
            L.ptr(self.ws), self.ws.numel(), C.byref(h)))
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

    def load(self, dest, weight, indptr, goff):
        """Copies one batch into the recorded buffers (stream ordered)."""
        t = L.torch()
        d_dest, d_w, d_ip, _ = _graph_to_device(dest, weight, indptr, None)
        N, M = d_ip.numel() - 1, d_dest.numel()
        g = np.ascontiguousarray(np.asarray(goff, dtype=np.int64))
        G = g.size - 1
        if N < 1 or M < 1 or G < 1:
            raise ValueError("LifetimeCutGraph: empty batch")
        if N > self.n_cap or M > self.m_cap or G > self.g_cap:
            raise ValueError(f"LifetimeCutGraph: batch (N={N}, M={M}, G={G}) exceeds the "
                             f"capacities ({self.n_cap}, {self.m_cap}, {self.g_cap})")
        if N == self.n_cap and M < self.m_cap:
            raise ValueError("LifetimeCutGraph: padding entries need a padding vertex (N < n_cap)")
        if N < self.n_cap and G >= self.g_cap:
            raise ValueError("LifetimeCutGraph: padding vertices need a padding graph (G < g_cap)")
        if g[0] != 0 or g[-1] != N or np.any(np.diff(g) < 0):
            raise ValueError("LifetimeCutGraph: bad graph offsets")
        if d_w.dtype != self.weight.dtype:
            raise TypeError(f"LifetimeCutGraph: weights are {d_w.dtype}, recorded for "
                            f"{self.weight.dtype}")
        self.dest[:M].copy_(d_dest)
        self.dest[M:].fill_(self.n_cap - 1)
        self.weight[:M].copy_(d_w)
        self.weight[M:].copy_(d_w[:1].expand(self.m_cap - M))
        self.indptr[:N + 1].copy_(d_ip)
        self.indptr[N + 1:].fill_(M)
        self.indptr[self.n_cap:].fill_(self.m_cap)
        gf = np.full(self.g_cap + 1, self.n_cap, dtype=np.int64)
        gf[:G + 1] = g
        self.goff.copy_(t.from_numpy(gf))
        return N, M, G

    def launch(self):
        L.check(self._lib.eac_sl_lifetime_batched_graph_launch(self._h, L.stream_ptr()))

    def run(self, dest, weight, indptr, goff):
        """Same results as lifetime_cut_batched(dest, weight, indptr, goff,
        rule, max_dist, n_clusters) for the recorded settings."""
        N, M, G = self.load(dest, weight, indptr, goff)
        self.launch()
        s = self.summary[:5].cpu().numpy()
        if s[4] & 16 and self.rounds < self.bound_rounds:
            self._capture(0)  # a batch needing more rounds than recorded
            self.launch()
            s = self.summary[:5].cpu().numpy()
        if s[4]:
            raise RuntimeError(f"LifetimeCutGraph: device error flags {int(s[4])}")
        if self.tune:
            want = min(int(s[2]) + self.margin, self.bound_rounds)
            self.tune = False
            if want < self.rounds:
                self._capture(want)
        gout = self.gout[:G]
        info = {"components": int(gout[:, 0].sum().item()), "forest_edges": int(s[1]),
                "rounds": int(s[2]), "kept_edges": int(s[3]), "recorded_rounds": self.rounds}
        return self.labels[:N], gout, info
