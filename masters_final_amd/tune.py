"""Autotuner for libeac_hip.so's launch and algorithm knobs (EAC_TUNE_*).

The knobs change speed only, never results: the Borůvka weight-level split
(light undirected edges per vertex x 10 below the first threshold, -1 = no
split; growth of that count per further level, -1 = one split only), the
number of rounds enqueued before the first read-back, the blocks per compute
unit of the edge-list filter (-1: one 1024-thread block per unit), the co-association LDS window (columns) and the
co-association member loads in flight per lane.
`autotune_mst` times a grid of candidate settings
on the caller's own graph with HIP events, checks that every candidate gives
the same forest, keeps the fastest, and caches the choice per device, weight
type and graph-shape bucket (JSON under $EAC_TUNE_CACHE, default
~/.cache/masters_final_amd/tune.json).
"""
from __future__ import annotations

import ctypes as C
import itertools
import json
import os
from pathlib import Path

from . import _lib as L

KNOBS = {"light_x10": 0, "first_batch": 1, "filter_bpcu": 2, "co_window": 3, "level_step": 4,
         "co_unroll": 5, "co_ranges": 6, "km_waves": 7}

DEFAULT_GRID = {
    "light_x10": (-1, 10, 15, 20, 30),
    "level_step": (-1, 0, 8),
    "first_batch": (0,),
    "filter_bpcu": (0, 4, -1),
}


def set_knobs(**values):
    lib = L.load_library()
    for name, v in values.items():
        L.check(lib.eac_tune_set(KNOBS[name], int(v)))


def get_knobs():
    lib = L.load_library()
    out = {}
    for name, k in KNOBS.items():
        v = C.c_int32()
        L.check(lib.eac_tune_get(k, C.byref(v)))
        out[name] = int(v.value)
    return out


def cache_path() -> Path:
    p = os.environ.get("EAC_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "masters_final_amd" / "tune.json"


def _load_cache():
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store_cache(c):
    try:
        path = cache_path()
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(c, indent=1, sort_keys=True))
    except OSError:
        pass  # read-only home: the choice still applies to this process


def mst_shape_key(n, m, dtype):
    t = L.torch()
    dev = t.cuda.get_device_name(t.cuda.current_device())
    return "mst|%s|%s|n2^%d|deg%d" % (dev, str(dtype).replace("torch.", ""),
                                      max(int(n), 1).bit_length(), round(m / max(n, 1)))


def autotune_mst(dest, weight, indptr, grid=None, reps=3, use_cache=True):
    """Picks the EAC_TUNE_* settings for Borůvka on this graph (device
    tensors: dest int32, weight, indptr int64[n+1]) and applies them.
    Returns dict(knobs, ms, candidates, cached)."""
    from .graph.mst import boruvka_mst
    t = L.torch()
    n, m = indptr.numel() - 1, dest.numel()
    key = mst_shape_key(n, m, weight.dtype)
    cache = _load_cache() if use_cache else {}
    if key in cache:
        set_knobs(**cache[key]["knobs"])
        return {"knobs": cache[key]["knobs"], "ms": cache[key]["ms"], "candidates": None,
                "cached": True}
    grid = grid or DEFAULT_GRID
    names = list(grid)
    results = []
    ref = None
    for combo in itertools.product(*(grid[k] for k in names)):
        knobs = dict(zip(names, combo))
        set_knobs(**knobs)
        r = boruvka_mst(dest, weight, indptr, None, want_ids=True, want_weights=True,
                        want_labels=True)  # warm-up (and result check)
        sig = (r["n_mst"], r["n_components"],
               int(r["weights"][: r["n_mst"]].to(t.int64).sum().item()))
        if ref is None:
            ref = sig
        elif sig != ref:
            raise RuntimeError(f"tuning knobs {knobs} changed the forest: {sig} vs {ref}")
        ev0, ev1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            boruvka_mst(dest, weight, indptr, None, want_ids=True, want_weights=True,
                        want_labels=True)
        ev1.record()
        ev1.synchronize()
        results.append((ev0.elapsed_time(ev1) / reps, knobs))
    results.sort(key=lambda x: x[0])
    best_ms, best = results[0]
    set_knobs(**best)
    cache[key] = {"knobs": best, "ms": best_ms}
    if use_cache:
        _store_cache(cache)
    return {"knobs": best, "ms": best_ms, "cached": False,
            "candidates": [{"ms": ms, **k} for ms, k in results]}


def autotune_coassoc(labels, ks, windows=(4096, 8192, 16384, 32768), unrolls=(4, 8),
                     use_cache=True):
    """Picks the co-association LDS window and the member loads in flight per
    lane for this ensemble (device labels [P][n], cluster counts ks): timed
    candidates, identical output checked (row offsets and the largest
    count). Returns dict(knobs, ms, ...)."""
    from .EAC.sparse import coassoc_device
    t = L.torch()
    P, n = labels.shape
    dev = t.cuda.get_device_name(t.cuda.current_device())
    key = "coassoc2|%s|n2^%d|P%d|k%d" % (dev, int(n).bit_length(), int(P),
                                          int(sum(int(k) for k in ks)).bit_length())
    cache = _load_cache() if use_cache else {}
    if key in cache:
        set_knobs(**cache[key]["knobs"])
        return {"knobs": cache[key]["knobs"], "ms": cache[key]["ms"], "candidates": None,
                "cached": True}
    results, ref = [], None
    for w, u in itertools.product(windows, unrolls):
        if w > 32768:
            continue
        set_knobs(co_window=w, co_unroll=u)
        ev0, ev1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        ev0.record()
        ip, ind, dat, mx = coassoc_device(labels, ks)
        ev1.record()
        ev1.synchronize()
        sig = (int(ip[-1].item()), int(mx), int(ip[n // 2].item()))
        if ref is None:
            ref = sig
        elif sig != ref:
            raise RuntimeError(f"co_window={w}, co_unroll={u} changed the co-association: "
                               f"{sig} vs {ref}")
        del ip, ind, dat
        results.append((ev0.elapsed_time(ev1), {"co_window": w, "co_unroll": u}))
    results.sort(key=lambda x: x[0])
    best_ms, best = results[0]
    set_knobs(**best)
    cache[key] = {"knobs": best, "ms": best_ms}
    if use_cache:
        _store_cache(cache)
    return {"knobs": best, "ms": best_ms, "cached": False,
            "candidates": [{"ms": ms, **k} for ms, k in results]}


def autotune_kmeans(X, centroids, k_list=None, waves=(0, 2, 3, 4), reps=3, use_cache=True):
    """Picks the waves per SIMD the split-BF16 labeling kernel is compiled
    for (km_waves; 0 = built-in) on this data shape: timed label passes,
    identical labels checked. Returns dict(knobs, ms, cached, candidates)."""
    from .cluster.KMeans import label_device
    t = L.torch()
    n, d = X.shape
    ks = list(k_list) if k_list is not None else [int(centroids.shape[0])]
    dev = t.cuda.get_device_name(t.cuda.current_device())
    key = "kmeans|%s|n2^%d|d%d|k2^%d|P%d" % (dev, int(n).bit_length(), int(d),
                                             int(sum(ks)).bit_length(), len(ks))
    cache = _load_cache() if use_cache else {}
    if key in cache:
        set_knobs(**cache[key]["knobs"])
        return {"knobs": cache[key]["knobs"], "ms": cache[key]["ms"], "candidates": None,
                "cached": True}
    results, ref = [], None
    for w in waves:
        set_knobs(km_waves=w)
        lab, _, _ = label_device(X, centroids, ks, precision="bf16x3")  # warm
        ev0, ev1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            lab, _, _ = label_device(X, centroids, ks, precision="bf16x3")
        ev1.record()
        ev1.synchronize()
        if ref is None:
            ref = lab
        elif not t.equal(lab, ref):
            raise RuntimeError(f"km_waves={w} changed the labels")
        results.append((ev0.elapsed_time(ev1) / reps, {"km_waves": w}))
    results.sort(key=lambda x: x[0])
    best_ms, best = results[0]
    set_knobs(**best)
    cache[key] = {"knobs": best, "ms": best_ms}
    if use_cache:
        _store_cache(cache)
    return {"knobs": best, "ms": best_ms, "cached": False,
            "candidates": [{"ms": ms, **k} for ms, k in results]}
