"""numpy wrappers around liboracle.so (the C restatement) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

_P = C.c_void_p
_I = C.c_int64
_D = C.c_double


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            from . import build as _b
            _b.build()
        L = C.CDLL(str(path))
        sig = {
            "orc_exprefixsum_i64": (_I, [_P, _P, _I, _I]),
            "orc_exprefixsum_single_i64": (_I, [_P, _I, _I]),
            "orc_boruvka_minho_seq": (_I, [_P, _P, _P, _P, _I, _I, _P, _P]),
            "orc_connected_comps_seq": (None, [_P, _P, _P, _P, _I, _I, _P]),
            "orc_kruskal_canonical": (_I, [_P, _P, _P, _I, _I, _P, _P]),
            "orc_components": (_I, [_P, _P, _I, _I, _P]),
            "orc_canonical_relabel": (_I, [_P, _I, _I, _P]),
            "orc_graph_from_edges": (_I, [_P, _P, _P, _P, _I, _P, _I, _P, _P, _P, _P]),
            "orc_sp_sl_lifetime": (C.c_int, [_P, _P, _P, _I, _I, _D, _I, _P, _P, _P, _P, _P]),
            "orc_sl_mst_lifetime": (C.c_int, [_P, _P, _P, _I, _I, _D, _P, _P, _P]),
            "orc_coassoc_condensed": (_I, [_P, _P, _I, _I, _P, _P, _P, _I]),
            "orc_coassoc_capped": (C.c_int, [_P, _I, _I, C.c_int, _P, _I, _P, _P, _P, _P, _P,
                                             _P]),
            "orc_coassoc_full": (None, [_P, _I, _I, C.c_int, _P]),
            "orc_boruvka_omp": (_I, [_P, _P, _P, _I, _P, _P, _P]),
            "orc_sl_batched_omp": (None, [_P, _P, _P, _P, _I, _P, _P, _P]),
            "orc_omp_threads": (C.c_int, []),
            "orc_argsort1d": (None, [_P, _I, _P]),
            "orc_slink": (None, [_P, _I, _P]),
            "orc_labels_from_Z": (None, [_P, _I, _I, _P]),
            "orc_full_sl_lifetime": (C.c_int, [_P, _I, _I, _P, _P, _P, _P]),
            "orc_kmeans_label_f64": (None, [_P, _P, _I, _I, _I, _P, _P]),
            "orc_arg_k_select": (None, [_P, _I, _I, _P]),
            "orc_recompute_centroids": (None, [_P, _P, _P, _I, _I, _I, C.c_int, _P]),
            "orc_kmeans_fit": (_I, [_P, _P, _I, _I, _I, _I, _D, C.c_int, _P, _P, _P, _P]),
            "orc_sum_f32": (C.c_float, [_P, _I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _a(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


def _p(a):
    return a.ctypes.data_as(_P)


# ------------------------------------------------------------------- scans
def exprefixsum(a, init=0):
    a = _a(a, np.int64)
    out = np.empty_like(a)
    tot = lib().orc_exprefixsum_i64(_p(a), _p(out), a.size, int(init))
    return out, int(tot)


def exprefixsum_single(a, init=0):
    a = _a(a, np.int64).copy()
    tot = lib().orc_exprefixsum_single_i64(_p(a), a.size, int(init))
    return a, int(tot)


# ----------------------------------------------------------------- graphs
def outdegree_from_firstedge(fe, n_edges):
    """build.py:275-280 / test_boruvka.py:37-42."""
    fe = np.asarray(fe)
    od = np.empty_like(fe)
    od[:-1] = fe[1:] - fe[:-1]
    od[-1] = n_edges - fe[-1]
    return od


def boruvka_minho_seq(dest, weight, fe, od):
    dest, fe, od = _a(dest, np.int64), _a(fe, np.int64), _a(od, np.int64)
    w = _a(weight, np.float64)
    nv = fe.size
    mst = np.zeros(max(nv - 1, 1), np.int64)
    comps = np.zeros(nv, np.int64)
    nm = lib().orc_boruvka_minho_seq(_p(dest), _p(w), _p(fe), _p(od), nv, dest.size,
                                     _p(mst), _p(comps))
    return mst[:max(nv - 1, 0)], int(nm), comps


def boruvka_omp(indptr, dest, weight):
    """All-core CPU Borůvka (OpenMP) of a symmetric int32-weight CSR: the
    canonical forest as (lo, hi) pairs, its weight, and the threads used."""
    ip = _a(indptr, np.int64)
    d = _a(dest, np.int32)
    w = _a(weight, np.int32)
    n = ip.size - 1
    lo = np.zeros(max(n, 1), np.int64)
    hi = np.zeros(max(n, 1), np.int64)
    tw = C.c_double()
    nf = lib().orc_boruvka_omp(_p(ip), _p(d), _p(w), n, _p(lo), _p(hi), C.byref(tw))
    if nf < 0:
        raise ValueError("weight range + ids exceed 64 bits")
    return lo[:nf], hi[:nf], int(tw.value), int(lib().orc_omp_threads())


def sl_batched_omp(indptr, dest, weight, goff):
    """All-core CPU cut of every graph of a disjoint union (OpenMP, largest
    graphs first): (components per graph, 0 where the reference returns empty
    labels, -1 where it raises; first-occurrence labels per graph; threads)."""
    ip = _a(indptr, np.int64)
    d = _a(dest, np.int32)
    w = _a(weight, np.int32)
    go = _a(goff, np.int64)
    G = go.size - 1
    order = np.argsort(-np.diff(go), kind="stable").astype(np.int64)
    nc = np.zeros(max(G, 1), np.int64)
    lab = np.zeros(max(int(go[-1]), 1), np.int64)
    lib().orc_sl_batched_omp(_p(ip), _p(d), _p(w), _p(go), G, _p(order), _p(nc), _p(lab))
    return nc[:G], lab[: int(go[-1])], int(lib().orc_omp_threads())


def connected_comps_seq(dest, weight, fe, od):
    dest, fe, od = _a(dest, np.int64), _a(fe, np.int64), _a(od, np.int64)
    w = _a(weight, np.float64)
    lab = np.zeros(fe.size, np.int64)
    lib().orc_connected_comps_seq(_p(dest), _p(w), _p(fe), _p(od), fe.size, dest.size, _p(lab))
    return lab


def csr_edges(dest, weight, fe, od, symmetric=True):
    """Undirected edge triples of a CSR: entries with dest > row (symmetric)
    or every non-loop entry (symmetric=False)."""
    fe, od, dest = np.asarray(fe, np.int64), np.asarray(od, np.int64), np.asarray(dest, np.int64)
    rows = np.repeat(np.arange(fe.size, dtype=np.int64), od)
    pos = np.concatenate([np.arange(f, f + o, dtype=np.int64) for f, o in zip(fe, od)]) \
        if fe.size and od.sum() else np.zeros(0, np.int64)
    cols = dest[pos]
    sel = cols > rows if symmetric else cols != rows
    return rows[sel], cols[sel], np.asarray(weight, np.float64)[pos][sel], pos[sel]


def kruskal_canonical(u, v, w, nv):
    u, v, w = _a(u, np.int64), _a(v, np.int64), _a(w, np.float64)
    inf = np.zeros(max(u.size, 1), np.uint8)
    order = np.zeros(max(u.size, 1), np.int64)
    nf = lib().orc_kruskal_canonical(_p(u), _p(v), _p(w), u.size, nv, _p(inf), _p(order))
    return inf[:u.size].astype(bool), order[:nf]


def components(u, v, nv):
    u, v = _a(u, np.int64), _a(v, np.int64)
    lab = np.zeros(max(nv, 1), np.int64)
    nc = lib().orc_components(_p(u), _p(v), u.size, nv, _p(lab))
    return int(nc), lab[:nv]


def canonical_relabel(labels):
    lab = _a(labels, np.int64)
    nv = int(lab.max()) + 1 if lab.size else 0
    out = np.zeros(max(lab.size, 1), np.int64)
    c = lib().orc_canonical_relabel(_p(lab), lab.size, nv, _p(out))
    return out[:lab.size], int(c)


def graph_from_edges(dest, weight, fe, od, edges):
    dest, fe, od, edges = (_a(dest, np.int64), _a(fe, np.int64), _a(od, np.int64),
                           _a(edges, np.int64))
    w = _a(weight, np.float64)
    nv = fe.size
    nod = np.zeros(nv, np.int64)
    nfe = np.zeros(nv, np.int64)
    ndest = np.zeros(max(2 * edges.size, 1), np.int64)
    nw = np.zeros(max(2 * edges.size, 1), np.float64)
    rc = lib().orc_graph_from_edges(_p(dest), _p(w), _p(fe), _p(od), nv, _p(edges), edges.size,
                                    _p(nod), _p(nfe), _p(ndest), _p(nw))
    if rc != 0:
        return None
    return nod, nfe, ndest[:2 * edges.size], nw[:2 * edges.size]


def sp_sl_lifetime(u, v, w, nv, max_dist=None, n_clusters=0):
    """eac.py:199-267 -> (n_comps, labels) with labels == [] when the
    reference returns an empty array."""
    u, v, w = _a(u, np.int64), _a(v, np.int64), _a(w, np.float64)
    lab = np.zeros(max(nv, 1), np.int64)
    nc, empty, th, ncs = C.c_int64(), C.c_int64(), C.c_double(), C.c_int64()
    rc = lib().orc_sp_sl_lifetime(_p(u), _p(v), _p(w), u.size, nv,
                                  -1.0 if max_dist is None else float(max_dist),
                                  int(n_clusters), _p(lab), C.byref(nc), C.byref(empty),
                                  C.byref(th), C.byref(ncs))
    if rc != 0:
        raise ValueError("sp_sl_lifetime: the reference raises ValueError here")
    if empty.value:
        return 1, np.zeros(0, np.int64)
    return int(nc.value), lab[:nv]


def sl_mst_lifetime(u, v, w, nv, disconnect_weight=None):
    u, v, w = _a(u, np.int64), _a(v, np.int64), _a(w, np.float64)
    lab = np.zeros(max(nv, 1), np.int64)
    nc, kept = C.c_int64(), C.c_int64()
    rc = lib().orc_sl_mst_lifetime(_p(u), _p(v), _p(w), u.size, nv,
                                   -1.0 if disconnect_weight is None else float(disconnect_weight),
                                   _p(lab), C.byref(nc), C.byref(kept))
    if rc != 0:
        raise ValueError("sl_mst_lifetime: the reference raises ValueError here")
    return lab[:nv], int(nc.value), int(kept.value)


def coassoc_condensed(labels, kp):
    """labels: (P, n) ints in [0, k_p), or < 0 for a point in no cluster of
    partition p. Returns upper-triangle CSR (indptr, indices, counts)."""
    L = _a(labels, np.int64)
    kp = _a(kp, np.int64)
    P, n = L.shape
    cap = 1  # upper bound: all within-cluster pairs of all partitions
    for p in range(P):
        c = np.bincount(L[p][L[p] >= 0], minlength=1).astype(np.int64)
        cap += int((c * (c - 1) // 2).sum())
    cap = min(cap, n * (n - 1) // 2 + 1)
    indptr = np.zeros(n + 1, np.int64)
    ind = np.zeros(cap, np.int64)
    cnt = np.zeros(cap, np.int64)
    nnz = lib().orc_coassoc_condensed(_p(L), _p(kp), P, n, _p(indptr), _p(ind), _p(cnt), cap)
    assert nnz >= 0
    return indptr, ind[:nnz], cnt[:nnz]


def coassoc_capped(labels, cap_off, full_rows=False, max_assocs=0):
    """The reference's capacity-capped surgical build (orc_coassoc_capped).
    labels (P, n), -1 = not listed; cap_off int64 [n+1] capacity offsets.
    Returns (indptr, indices, counts, discarded); raises ValueError (with the
    row) when the first partition overflows a row."""
    L = _a(labels, np.int64)
    off = _a(cap_off, np.int64)
    P, n = L.shape
    tot = max(int(off[-1]), 1)
    indptr = np.zeros(n + 1, np.int64)
    ind = np.zeros(tot, np.int64)
    cnt = np.zeros(tot, np.int64)
    nnz, disc, bad = C.c_int64(), C.c_int64(), C.c_int64()
    rc = lib().orc_coassoc_capped(_p(L), P, n, 1 if full_rows else 0, _p(off), int(max_assocs),
                                  _p(indptr), _p(ind), _p(cnt), C.byref(nnz), C.byref(disc),
                                  C.byref(bad))
    if rc != 0:
        raise ValueError(f"first partition overflows row {bad.value}")
    return indptr, ind[:nnz.value], cnt[:nnz.value], int(disc.value)


# ------------------------------------------------- dense EAC + single link
def coassoc_full(labels, condensed=True):
    """EAC_FULL counts (full.py:141-166): condensed n(n-1)/2 or n x n int64."""
    L = _a(labels, np.int64)
    P, n = L.shape
    out = np.zeros(n * (n - 1) // 2 if condensed else n * n, np.int64)
    lib().orc_coassoc_full(_p(L), P, n, 1 if condensed else 0, _p(out))
    return out if condensed else out.reshape(n, n)


def argsort1d(values):
    """scipy_numba's argsort1D (quicksort, first-element pivot)."""
    v = _a(values, np.float64)
    out = np.zeros(v.size, np.int64)
    lib().orc_argsort1d(_p(v), v.size, _p(out))
    return out


def slink(dists, n):
    """SLINK + from_pointer_representation: Z (n-1, 4) float64."""
    d = _a(dists, np.float64)
    Z = np.zeros((max(n - 1, 1), 4), np.float64)
    lib().orc_slink(_p(d), n, _p(Z))
    return Z[: n - 1]


def labels_from_Z(Z, n_clusters):
    Z = _a(Z, np.float64)
    n = Z.shape[0] + 1
    lab = np.zeros(n, np.int64)
    lib().orc_labels_from_Z(_p(Z), n, int(n_clusters), _p(lab))
    return lab


def full_sl_lifetime(dists, n, n_clusters=0):
    """eac.py:337-364 on condensed dissimilarities: (nc, labels, Z); labels
    empty for one cluster; ValueError where numpy raises."""
    d = _a(dists, np.float64)
    Z = np.zeros((max(n - 1, 1), 4), np.float64)
    lab = np.zeros(n, np.int64)
    nc, empty = C.c_int64(), C.c_int64()
    rc = lib().orc_full_sl_lifetime(_p(d), n, int(n_clusters), _p(Z), _p(lab), C.byref(nc),
                                    C.byref(empty))
    if rc != 0:
        raise ValueError("lifetime rule: no lifetime close to zero (numpy raises)")
    return int(nc.value), (np.zeros(0, np.int64) if empty.value else lab), Z[: n - 1]


# ---------------------------------------------------------------- K-Means
def kmeans_label_f64(X, Cc):
    X, Cc = _a(X, np.float32), _a(Cc, np.float32)
    n, d = X.shape
    k = Cc.shape[0]
    lab = np.zeros(n, np.int32)
    dist = np.zeros(n, np.float64)
    lib().orc_kmeans_label_f64(_p(X), _p(Cc), n, d, k, _p(lab), _p(dist))
    return lab, dist


def arg_k_select(ary, k):
    a = _a(ary, np.float32)
    out = np.zeros(max(k, 1), np.int64)
    lib().orc_arg_k_select(_p(a), a.size, k, _p(out))
    return out[:k]


def recompute_centroids(X, labels, dists, k, reseed_mode=0):
    X = _a(X, np.float32)
    lab = _a(labels, np.int32)
    dist = _a(dists, np.float32)
    n, d = X.shape
    out = np.zeros((k, d), np.float32)
    lib().orc_recompute_centroids(_p(X), _p(lab), _p(dist), n, d, k, reseed_mode, _p(out))
    return out


def sum_f32(a):
    """numpy's float32 `a.sum()` (pairwise summation), restated in C."""
    a = _a(a, np.float32)
    return np.float32(lib().orc_sum_f32(_p(a), a.size))


def kmeans_fit(X, init, max_iter=3, tol=1e-4, reseed_mode=0, return_inertia=False):
    X, init = _a(X, np.float32), _a(init, np.float32)
    n, d = X.shape
    k = init.shape[0]
    lab = np.zeros(n, np.int32)
    dist = np.zeros(n, np.float32)
    cen = np.zeros((k, d), np.float32)
    inertia = C.c_float()
    it = lib().orc_kmeans_fit(_p(X), _p(init), n, d, k, max_iter, tol, reseed_mode, _p(lab),
                              _p(dist), _p(cen), C.byref(inertia))
    if return_inertia:
        return lab, dist, cen, int(it), np.float32(inertia.value)
    return lab, dist, cen, int(it)
