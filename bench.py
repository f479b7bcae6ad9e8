"""Benchmarks of the EAC hot path on one MI355X.

Default workload (the one the headline metric is quoted on): Boruvka minimum
spanning forest + connected components of a 1,000,000-vertex symmetrized 3-D
k-NN graph (k=20, int32 weights uniform in [1,100], 10% of the vertices in 4
separate islands, seed 1). One step = one call of the forest + component-label
path, inputs resident in HBM; value = directed CSR entries / step time
(Medges/s). The CPU baseline is the single-threaded C restatement of the
reference's boruvka_minho_seq on the same graph, with an all-core OpenMP
Borůvka line beside it. The same JSON line carries the end-to-end EAC
comparison on the 10,000-point workload (batched ensemble vs plain per-call
FP32 path, exact labels) and, unless --no-extras, every other workload as an
extra key with its own roofline and CPU baseline: the same 1M-vertex graph at
mean degree ~40 (k=37), the 4,194,304 x 64 x 2048 labeling, the 500,000-point
256-partition ensemble + co-association, and the 512-graph batched cut.

Other workloads (--workload): eac, kmeans, coassoc, sl_batched, accuracy.
Each prints one JSON line with the same top-level keys.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "EAC pts/s + Boruvka MST Medges/s: batched/BF16 vs plain FP32 GPU, exact labels"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3
MST_SLOTS = {0: "k_round0 (CSR stream)", 1: "k_round (edge-list pass)",
             2: "k_hook (pick-tree walks)", 4: "forest output + labels"}


def profile_traffic(kernel, summary="boruvka_mst_knn1m_hbm_traffic.md"):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (scaled by a copy-kernel calibration) and committed under
    profiles/; None when no summary covers it."""
    import re
    path = Path(__file__).resolve().parent / "profiles" / summary
    if not path.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch %s: (\d+)" % re.escape(kernel), path.read_text())
    return int(m.group(1)) if m else None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    """The host the CPU baselines ran on: logical CPUs, the CPUs this process
    may use, OMP_NUM_THREADS and the CPU model (/proc/cpuinfo)."""
    import os
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def read_prof(lib, L, slots):
    prof = {}
    for slot in slots:
        tm, nl, nb = C.c_double(), C.c_int64(), C.c_int64()
        L.check(lib.eac_profile_read(slot, C.byref(tm), C.byref(nl), C.byref(nb)))
        prof[slot] = (tm.value, nl.value, nb.value)
    return prof


def base_line(args, value, unit, ms, dtype, config, higher=True):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": higher,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": config}


def timed(fn, steps, warmup, torch):
    for _ in range(warmup):
        r = fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = fn()
    torch.cuda.synchronize()
    return r, (time.perf_counter() - t0) * 1e3 / steps


# ------------------------------------------------------------------ EAC e2e
def eac_compare(n=10_000, P=100, kmin=50, kmax=100, iters=3, seed=0, reps=3, cpu=True):
    """Batched+BF16 EAC vs plain per-call FP32 EAC on the 2-D 6-GMM workload.
    k ~ U-int[kmin, kmax] inclusive, initial centroids = distinct random points."""
    import numpy as np
    import torch

    from masters_final_amd import synth
    from masters_final_amd.cluster.KMeans import fit_batched
    from masters_final_amd.EAC.eac import eac_pipeline, sp_sl_lifetime_device
    from masters_final_amd.EAC.sparse import coassoc_device

    X, _ = synth.gmm_points(n, 2, 6, seed, mean_scale=10.0, std=1.0)
    rng = np.random.default_rng(seed)
    ks = rng.integers(kmin, kmax + 1, P)
    init = np.concatenate([X[rng.choice(n, int(k), replace=False)] for k in ks])
    dX = torch.from_numpy(X).cuda()
    dinit = torch.from_numpy(init).cuda()
    off = np.concatenate([[0], np.cumsum(ks)])

    def batched():
        return eac_pipeline(dX, ks, iters=iters, precision="bf16x3", init=dinit)

    def plain():
        labs = []
        for p in range(P):  # one K-Means call per partition, FP32 labeling
            r = fit_batched(dX, dinit[off[p]:off[p + 1]], [int(ks[p])], max_iter=iters,
                            precision="fp32")
            labs.append(r["labels"])
        labels = torch.cat(labs)
        ip, ind, dat, mx = coassoc_device(labels, ks, w_offset="max+1")
        nc, lab, _ = sp_sl_lifetime_device((ip, ind, dat), P)
        return {"n_clusters": nc, "labels": lab, "ensemble_labels": labels}

    rb, tb = timed(batched, reps, 1, torch)
    # the same pipeline with the ensemble and the co-association replayed
    # from one HIP graph (recorded once, outside the timing), then the cut
    from masters_final_amd.EAC.eac import EnsembleCoassocGraph, eac_pipeline_graph
    gpipe = EnsembleCoassocGraph(dX, ks, iters=iters, precision="bf16x3")
    rg, tg = timed(lambda: eac_pipeline_graph(gpipe, dinit), reps, 1, torch)
    # the whole path recorded: the cut as a second HIP graph sized on one
    # eager pass (capacity nnz x 1.25, rounds tuned on the first replay)
    from masters_final_amd.EAC.eac import EACPipelineGraph
    rec = EACPipelineGraph(dX, ks, dinit, iters=iters, precision="bf16x3")
    rr, tr = timed(lambda: rec.run(dinit), reps, 1, torch)
    rp, tp = timed(plain, reps, 1, torch)
    out = {"n": n, "partitions": P, "k_range": [kmin, kmax], "iters": iters,
           "batched_labeling": "FP64 direct difference: d = 2 is below the d >= 16 where the "
                               "split-BF16 MFMA path applies, so precision 'bf16x3' runs the "
                               "FP64 labeler here; the gain is from batching the partitions",
           "plain_labeling": "FP32 direct difference, one K-Means call per partition",
           "batched_bf16_ms": tb, "plain_fp32_ms": tp, "speedup": tp / tb,
           "graph_replay_ms": tg, "graph_speedup_vs_plain": tp / tg,
           "graph_replay_scope": "ensemble (3 labelings) + one-pass co-association replayed "
                                 "from one HIP graph, then one read-back of the matrix size "
                                 "and the forest + cut launched eagerly",
           "recorded_ms": tr, "recorded_speedup_vs_plain": tp / tr,
           "recorded_scope": "ensemble + co-association graph, then the forest and the cut "
                             "as a recorded batched-cut graph; one read-back at the end",
           "recorded_rounds": rec.rounds,
           "batched_pts_per_s": n / (tb * 1e-3), "plain_pts_per_s": n / (tp * 1e-3),
           "n_clusters": rb["n_clusters"], "nnz": rb["nnz"], "rescored": rb["n_rescored"]}
    lb = rb["labels"].cpu().numpy() if rb["labels"] is not None else np.zeros(0)
    lp = rp["labels"].cpu().numpy() if rp["labels"] is not None else np.zeros(0)
    out["plain_equals_batched"] = bool(np.array_equal(lb, lp))
    lg = rg["labels"].cpu().numpy() if rg["labels"] is not None else np.zeros(0)
    out["graph_equals_batched"] = bool(np.array_equal(lb, lg) and
                                       rg["n_clusters"] == rb["n_clusters"])
    lr = rr["labels"].cpu().numpy() if rr["labels"] is not None else np.zeros(0)
    out["recorded_equals_batched"] = bool(np.array_equal(lb, lr) and
                                          rr["n_clusters"] == rb["n_clusters"])
    if cpu:
        from oracle import oracle as O
        c0 = time.perf_counter()
        labs = np.stack([O.kmeans_fit(X, init[off[p]:off[p + 1]], max_iter=iters,
                                      reseed_mode=1)[0] for p in range(P)])
        ip, ind, cnt = O.coassoc_condensed(labs, ks)
        w = (cnt.max() + 1 - cnt).astype(np.float64)
        rows = np.repeat(np.arange(n), np.diff(ip))
        nc, lab = O.sp_sl_lifetime(rows, ind, w, n, max_dist=P)
        c1 = time.perf_counter()
        out["oracle_equal"] = bool(np.array_equal(rb["ensemble_labels"].cpu().numpy(), labs)
                                   and nc == rb["n_clusters"] and np.array_equal(lb, lab))
        out["cpu_seconds"] = c1 - c0
        out["cpu_pts_per_s"] = n / (c1 - c0)
    return out


# --------------------------------------------------------------- workloads
def run_mst(args):
    out = mst_case(args, args.knn)
    if not args.no_eac:
        log("end-to-end EAC: batched+BF16 vs plain per-call FP32 (n=10,000, P=100)")
        # 10 host-timed repetitions: with fewer, launch and first-call
        # costs of a ~2 ms pipeline show in the mean
        # the library's built-in Borůvka choices: the knobs tuned on the
        # 1M-vertex graph fit that shape only (on the pipeline's 10K-vertex,
        # 4.8M-entry co-association graph one weight level costs +60%)
        from masters_final_amd.tune import get_knobs, set_knobs
        saved = get_knobs()
        set_knobs(light_x10=0, level_step=0, first_batch=0, filter_bpcu=0)
        try:
            out["eac_small"] = eac_compare(cpu=not args.no_cpu, reps=10)
        finally:
            set_knobs(**saved)
        log(f"EAC speedup {out['eac_small']['speedup']:.2f}x")
    if not args.no_extras:
        out.update(extras(args))
    out["host"] = host_info()
    return out


_LINE_KEYS = ("metric", "n_gpus", "steps", "warmup", "higher_is_better", "scaling",
              "vs_baseline", "data")


def _compact(line, steps, warmup):
    """An extra workload's line without the keys the headline line holds."""
    d = {k: v for k, v in line.items() if k not in _LINE_KEYS}
    d["steps"], d["warmup"] = steps, warmup
    return d


def extras(args):
    """The other benchmark workloads, each with its own roofline and CPU
    baseline, as extra keys of the default line: the 1M-vertex graph at mean
    degree ~40 (k=37), the 4,194,304 x 64 x 2048 labeling, the 500,000-point
    256-partition ensemble + co-association, the 512-graph batched cut."""
    import argparse as _ap
    ns = _ap.Namespace(**vars(args))
    ns.n = 1_000_000  # each workload's own default size
    ns.steps, ns.warmup = min(args.steps, 5), 1
    ns.knobs = ""  # every extra workload tunes its own knobs
    out = {}
    for key, fn in (("mst_knn_deg40", lambda a: mst_case(a, 37)),
                    ("kmeans_4m_d64_k2048", run_kmeans),
                    ("coassoc_500k_p256", run_coassoc),
                    ("sl_batched_512_knn", run_sl_batched)):
        log(f"extra workload {key}")
        t0 = time.perf_counter()
        out[key] = _compact(fn(ns), ns.steps, ns.warmup)
        log(f"extra workload {key} done in {time.perf_counter() - t0:.1f}s")
    return out


def mst_case(args, knn):
    import numpy as np
    import torch

    from masters_final_amd import _lib as L
    from masters_final_amd import synth
    from masters_final_amd.graph.mst import boruvka_mst

    lib = L.load_library()
    log(f"generating k-NN graph n={args.n} k={knn}")
    t0 = time.perf_counter()
    g = synth.knn_graph(n=args.n, k=knn, seed=1)
    n, m = g["n"], int(g["dest"].size)
    log(f"graph: {n} vertices, {m} directed entries ({m / n:.2f}/vertex), "
        f"{time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda", 0)
    d_dest = torch.from_numpy(g["dest"]).to(dev)
    d_w = torch.from_numpy(g["weight"]).to(dev)
    d_ip = torch.from_numpy(g["indptr"]).to(dev)

    def step():
        return boruvka_mst(d_dest, d_w, d_ip, None, want_ids=True, want_weights=True,
                           want_labels=True)

    # knobs (light/heavy split, first round batch, filter blocks per CU) from
    # the autotuner: timed candidates on this graph, same forest for each
    from masters_final_amd.tune import autotune_mst, set_knobs
    if args.knobs:  # fixed settings (profiling runs): no candidate timing
        knobs = {k: int(v) for k, v in (kv.split("=") for kv in args.knobs.split(","))}
        set_knobs(**knobs)
        tuning = {"knobs": knobs, "ms": None, "cached": None, "source": "--knobs"}
    else:
        tuning = autotune_mst(d_dest, d_w, d_ip, use_cache=not args.retune)
    log(f"tuning: {tuning['knobs']} (cached={tuning['cached']})")

    for _ in range(max(args.warmup, 1)):  # at least one call: the run's shape
        r = step()
    torch.cuda.synchronize()
    log(f"warmup done: forest {r['n_mst']} edges, {r['n_components']} components, "
        f"{r['rounds']} rounds")
    r, ms = timed(step, args.steps, 0, torch)  # value: library profiler off
    # kernel durations from a separate pass with the library's HIP-event
    # profiler on (its events add gaps between the bracketed kernels)
    L.check(lib.eac_profile_enable(1))
    timed(step, args.steps, 0, torch)
    prof = read_prof(lib, L, MST_SLOTS)
    L.check(lib.eac_profile_enable(0))
    dom = max(prof, key=lambda s: prof[s][0])
    tm, nl, nb = prof[dom]
    avg_ms = tm / max(nl, 1)
    bpl = nb / max(nl, 1)
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    log("kernel time per step: " + ", ".join(
        f"{MST_SLOTS[s]} {prof[s][0] / args.steps:.3f} ms" for s in prof))
    total_w = int(r["weights"][: r["n_mst"]].to(torch.int64).sum().item())
    labels_crc = int(torch.sum(r["labels"].to(torch.int64) *
                               torch.arange(n, device=dev, dtype=torch.int64) % 1000003).item())
    cpu = cpu_all = None
    if not args.no_cpu:
        from oracle import oracle as O
        log("timing the CPU restatement of boruvka_minho_seq (1 thread)")
        c0 = time.perf_counter()
        mst, nm, _ = O.boruvka_minho_seq(g["dest"], g["weight"], g["indptr"][:-1],
                                         np.diff(g["indptr"]))
        c1 = time.perf_counter()
        cpu_w = int(g["weight"][mst[:nm]].astype(np.int64).sum())
        if cpu_w != total_w or nm != r["n_mst"]:
            raise SystemExit(f"CPU/GPU forest mismatch: weight {cpu_w} vs {total_w}, "
                             f"edges {nm} vs {r['n_mst']}")
        cpu = {"value": m / (c1 - c0) / 1e6, "unit": "Medges/s", "cores": 1, "kind": "port",
               "sample": f"full graph ({n} vertices, {m} entries), one run of the C "
                         "restatement of mst.py:143-407 boruvka_minho_seq",
               "seconds": c1 - c0}
        log(f"CPU baseline {c1 - c0:.2f}s")
        # all cores: OpenMP Borůvka over the same CSR (the canonical forest)
        O.boruvka_omp(g["indptr"], g["dest"], g["weight"])  # warm-up (threads, pages)
        c0 = time.perf_counter()
        lo, hi, omp_w, threads = O.boruvka_omp(g["indptr"], g["dest"], g["weight"])
        c1 = time.perf_counter()
        if omp_w != total_w or lo.size != r["n_mst"]:
            raise SystemExit(f"all-core CPU/GPU forest mismatch: weight {omp_w} vs {total_w}")
        cpu_all = {"value": m / (c1 - c0) / 1e6, "unit": "Medges/s", "cores": threads,
                   "kind": "port", "sample": "full graph, OpenMP Borůvka with the same "
                   "(weight, min id, max id) order (oracle/eac_oracle.c orc_boruvka_omp)",
                   "seconds": c1 - c0}
        log(f"all-core CPU baseline {c1 - c0:.2f}s on {threads} threads")
    out = base_line(args, m / (ms * 1e-3) / 1e6, "Medges/s", ms, "int32",
                    {"workload": "boruvka_mst_cc_knn", "n_vertices": n, "m_entries": m,
                     "mean_degree": m / n, "knn_k": knn, "weights": "int32 U[1,100]",
                     "islands": 4, "seed": 1,
                     "step": "forest ids+weights+first-occurrence component labels"})
    out["result"] = {"forest_edges": r["n_mst"], "forest_weight": total_w,
                     "components": r["n_components"], "rounds": r["rounds"],
                     "labels_checksum": labels_crc}
    from masters_final_amd.tune import get_knobs
    big = get_knobs()["filter_bpcu"] < 0  # the 1024-thread filter variant
    kname = {0: "k_round0<int, false>", 1: "k_round<1024>" if big else "k_round<256>"}.get(dom)
    out["roofline"] = {"bound": "hbm", "kernel": MST_SLOTS[dom], "achieved": achieved,
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                       # the committed counter summary covers the k=20 graph only
                       "traffic": profile_traffic(kname) if kname and knn == 20 else None,
                       "traffic_source": "profiles/boruvka_mst_knn1m_hbm_traffic.md",
                       "avg_launch_ms": avg_ms, "alg_bytes_per_launch": bpl}
    out["kernels_ms_per_step"] = {MST_SLOTS[s]: prof[s][0] / args.steps for s in prof}
    out["tuning"] = tuning
    out["cpu_baseline"] = cpu
    out["cpu_baseline_all_cores"] = cpu_all
    return out


def run_eac(args):
    import torch
    r = eac_compare(n=args.n if args.n != 1_000_000 else 10_000, cpu=not args.no_cpu,
                    reps=max(args.steps, 1))
    out = base_line(args, r["batched_pts_per_s"], "pts/s", r["batched_bf16_ms"],
                    "f64 labeling (d=2), int32 counts",
                    {"workload": "eac_e2e_gmm2d", "n": r["n"], "partitions": r["partitions"],
                     "k_range": r["k_range"], "iters": r["iters"]})
    out["result"] = r
    out["cpu_baseline"] = ({"value": r["cpu_pts_per_s"], "unit": "pts/s", "cores": 1,
                            "kind": "port", "sample": "whole workload: C restatement of "
                            "K_Means.fit x P, EAC_CSR build, sp_sl_lifetime"}
                           if "cpu_pts_per_s" in r else None)
    return out


def run_kmeans(args):
    import numpy as np
    import torch

    from masters_final_amd import _lib as L
    from masters_final_amd import synth
    from masters_final_amd.cluster.KMeans import label_device

    lib = L.load_library()
    n = args.n if args.n != 1_000_000 else 4_194_304
    d, k = 64, 2048
    log(f"generating GMM n={n} d={d} (2048 components)")
    X, _ = synth.gmm_points(n, d, 2048, 2, mean_std=2.0, std=1.0)
    rng = np.random.default_rng(2)
    Cc = X[rng.choice(n, k, replace=False)].copy()
    dX = torch.from_numpy(X).cuda()
    dC = torch.from_numpy(Cc).cuda()
    from masters_final_amd.tune import autotune_kmeans, set_knobs
    if args.knobs:  # fixed settings (profiling runs): no candidate timing
        knobs = {kk: int(v) for kk, v in (kv.split("=") for kv in args.knobs.split(","))}
        set_knobs(**knobs)
        tuning = {"knobs": knobs, "ms": None, "cached": None, "source": "--knobs"}
    else:
        tuning = autotune_kmeans(dX, dC, use_cache=not args.retune)
    log(f"tuning: {tuning['knobs']} ({tuning['ms']} ms, cached={tuning['cached']})")
    res = {}
    for prec in ("bf16x3", "fp32", "fp64"):
        L.check(lib.eac_profile_enable(1))
        (lab, dist, nres), ms = timed(lambda: label_device(dX, dC, precision=prec),
                                      args.steps, args.warmup, torch)
        prof = read_prof(lib, L, {6: "label", 7: "rescore"})
        L.check(lib.eac_profile_enable(0))
        tm, nl, fl = prof[6]
        res[prec] = {"ms": ms, "labels": lab, "rescored": nres,
                     "label_kernel_ms": tm / max(nl, 1),
                     "label_tflops": (fl / max(nl, 1)) / (tm / max(nl, 1) * 1e-3) / 1e12,
                     "rescore_ms": prof[7][0] / max(prof[7][1], 1)}
        log(f"{prec}: {ms:.2f} ms/pass, rescored {nres}")
    ref = res["fp64"]["labels"]
    agree = {p: int((res[p]["labels"] != ref).sum().item()) for p in ("bf16x3", "fp32")}
    from oracle import oracle as O
    samp = rng.choice(n, 2000, replace=False)
    c0 = time.perf_counter()
    want, _ = O.kmeans_label_f64(X[samp], Cc)
    c1 = time.perf_counter()
    orc_ok = bool(np.array_equal(res["bf16x3"]["labels"][0].cpu().numpy()[samp], want))
    b = res["bf16x3"]
    flops = 2.0 * n * k * d
    out = base_line(args, flops / (b["ms"] * 1e-3) / 1e12, "TFLOP/s (2nkd)", b["ms"],
                    "bf16x3->fp32, fp64 rescore",
                    {"workload": "kmeans_label_gmm", "n": n, "d": d, "k": k, "seed": 2})
    out["result"] = {"bf16x3_ms": b["ms"], "fp32_ms": res["fp32"]["ms"],
                     "fp64_ms": res["fp64"]["ms"], "speedup_vs_fp32": res["fp32"]["ms"] / b["ms"],
                     "rescored_points": b["rescored"], "rescored_frac": b["rescored"] / n,
                     "mismatch_vs_fp64_gpu": agree, "oracle_sample_equal": orc_ok}
    out["tuning"] = tuning
    out["roofline"] = {"bound": "mfma", "kernel": "k_label_bf16x3",
                       "achieved": b["label_tflops"], "peak": BF16_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": b["label_tflops"] / BF16_PEAK_TFLOPS,
                       "traffic": None, "avg_launch_ms": b["label_kernel_ms"],
                       "note": "algorithmic 2nkd; the split issues 3x that on MFMA"}
    out["cpu_baseline"] = {"value": 2000 * k * d * 2 / (c1 - c0) / 1e12, "unit": "TFLOP/s",
                           "cores": 1, "kind": "port",
                           "sample": "2000 points x 2048 centroids, FP64 oracle labeling"}
    return out


def co_window_py(n, knob):
    """The co-association LDS window the library picks for n points (the
    window knob, else 32768, at most the row range rounded up to 64)."""
    w = knob if knob and knob > 0 else 32768
    w = min(w, 32768, -(-n // 64) * 64)
    return max(64, w & ~1)


def run_coassoc(args):
    import numpy as np
    import torch

    from masters_final_amd import _lib as L
    from masters_final_amd import synth
    from masters_final_amd.EAC.sparse import coassoc_device
    from masters_final_amd.utils.partition import generate_ensemble_batched

    lib = L.load_library()
    n = args.n if args.n != 1_000_000 else 500_000
    P = args.partitions
    log(f"generating anisotropic GMM n={n} d=16, P={P}")
    X, _ = synth.gmm_points(n, 16, 32, 3, mean_scale=20.0, anisotropic=(0.2, 3.0))
    rng = np.random.default_rng(3)
    ks = rng.integers(32, 2049, P)
    dX = torch.from_numpy(X).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    labels, fit = generate_ensemble_batched(dX, ks, iters=3, seed=3, precision="bf16x3")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f"ensemble {t1 - t0:.2f}s, rescored {fit['n_rescored']}")
    # the same ensemble (same initial centroids) timed eagerly and replayed
    # from a HIP graph (eac_kmeans_fit_graph: no host read-back)
    from masters_final_amd.cluster.KMeans import BatchedFitGraph, fit_batched
    irng = np.random.default_rng(3)
    idx = np.concatenate([irng.choice(n, int(k), replace=False) for k in ks])
    dinit = dX.index_select(0, torch.from_numpy(idx).to(dX.device))
    _, ens_ms = timed(lambda: fit_batched(dX, dinit, list(ks), max_iter=3, precision="bf16x3"),
                      2, 1, torch)
    gfit = BatchedFitGraph(dX, ks, max_iter=3, precision="bf16x3")
    _, ens_graph_ms = timed(lambda: gfit.run(dinit), 3, 1, torch)
    gres = gfit.result()
    ens_graph_equal = bool(torch.equal(gres["labels"], labels))
    del gfit
    log(f"ensemble eager {ens_ms:.1f} ms, graph replay {ens_graph_ms:.1f} ms, "
        f"equal {ens_graph_equal}")
    from masters_final_amd.tune import autotune_coassoc, set_knobs
    if args.knobs:  # fixed settings (profiling runs): no candidate timing
        knobs = {k: int(v) for k, v in (kv.split("=") for kv in args.knobs.split(","))}
        set_knobs(**knobs)
        tuning = {"knobs": knobs, "ms": None, "cached": None, "source": "--knobs"}
    else:
        tuning = autotune_coassoc(labels, ks, use_cache=not args.retune)
    log(f"tuning: {tuning['knobs']} ({tuning['ms']} ms, cached={tuning['cached']})")
    L.check(lib.eac_profile_enable(1))
    info = {}
    (ip, ind, dat, mx), ms = timed(lambda: coassoc_device(labels, ks, info=info), args.steps,
                                   args.warmup, torch)
    prof = read_prof(lib, L, {8: "plan", 9: "count", 10: "fill", 11: "stage", 12: "compact"})
    L.check(lib.eac_profile_enable(0))
    log(f"build path: {info}")
    nnz = int(ind.numel())
    incr = 0
    lab_h = labels.cpu().numpy()
    for p in range(P):
        s = np.bincount(lab_h[p], minlength=int(ks[p])).astype(np.int64)
        incr += int((s * (s - 1) // 2).sum())
    out = base_line(args, incr / (ms * 1e-3) / 1e9, "Gpair-increments/s", ms, "int32",
                    {"workload": "coassoc_batched_ensemble", "n": n, "d": 16, "partitions": P,
                     "k_range": [32, 2048], "seed": 3})
    def per_launch(slot):
        return prof[slot][0] / max(prof[slot][1], 1) if prof[slot][1] else 0.0

    out["result"] = {"nnz_condensed": nnz, "max_count": mx, "pair_increments": incr,
                     "ensemble_seconds": t1 - t0, "rescored": fit["n_rescored"],
                     "ensemble_eager_ms": ens_ms, "ensemble_graph_replay_ms": ens_graph_ms,
                     "ensemble_graph_equal": ens_graph_equal,
                     "build_path": info.get("path"),
                     "sample_stride": info.get("sample_stride"),
                     "nnz_estimate": info.get("estimate"),
                     "plan_ms": per_launch(8), "count_ms": per_launch(9),
                     "fill_ms": per_launch(10), "stage_ms": per_launch(11),
                     "compact_ms": per_launch(12)}
    out["tuning"] = tuning
    whole = 4.0 * P * n + 8.0 * nnz + 8.0 * (n + 1)  # labels read, (col, count) + offsets
    traffic = None
    if info.get("path") == "one-pass":
        # staged pass: member lists read once (4 B per (partition, point)),
        # 4 B per staged entry, 8 B per (row, window) segment, 4 B per row
        win = co_window_py(n, tuning["knobs"].get("co_window", 0))
        nw, q = -(-n // win), n // win
        # segment slots: sum over rows of nw - (i + 1) // win
        slots = n * nw - (win * q * (q - 1) // 2 + q * (n - q * win + 1))
        kname, k_ms = "k_coassoc_rows<stage>", per_launch(11)
        if n == 500_000 and P == 256:  # the profiled workload
            traffic = profile_traffic(kname, "coassoc_ensemble_n500k_p256_counters.md")
        k_bytes = 4.0 * P * n + 4.0 * nnz + 8.0 * slots + 4.0 * n
        c_ms = per_launch(12)
        c_bytes = 12.0 * nnz + 8.0 * slots + 8.0 * (n + 1)  # entries in, (col, value) out
        out["compaction"] = {"kernel": "k_coassoc_compact", "avg_launch_ms": c_ms,
                             "alg_bytes_per_launch": c_bytes,
                             "achieved": c_bytes / (c_ms * 1e-3) / 1e9 if c_ms else None,
                             "frac": c_bytes / (c_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if c_ms
                             else None}
    else:
        kname, k_ms = "k_coassoc_rows<fill>", per_launch(10)
        k_bytes = prof[10][2] / max(prof[10][1], 1)
    ach = k_bytes / (k_ms * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "kernel": kname, "achieved": ach,
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                       "traffic": traffic,
                       "traffic_source": "profiles/coassoc_ensemble_n500k_p256_counters.md"
                       if traffic else None, "avg_launch_ms": k_ms,
                       "alg_bytes_per_launch": k_bytes,
                       "whole_build_alg_bytes": whole,
                       "whole_build_frac": whole / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    out["cpu_baseline"] = None
    if not args.no_cpu:
        # the C restatement of the reference's surgical build on a sample of
        # the points (all 256 partitions, labels kept), checked against the
        # GPU build of the same sample
        from oracle import oracle as O
        ns = 20_000
        samp = np.sort(np.random.default_rng(33).choice(n, ns, replace=False))
        Ls = np.ascontiguousarray(lab_h[:, samp])
        sinc = 0
        for p in range(P):
            s = np.bincount(Ls[p], minlength=int(ks[p])).astype(np.int64)
            sinc += int((s * (s - 1) // 2).sum())
        c0 = time.perf_counter()
        sip, sind, scnt = O.coassoc_condensed(Ls, ks)
        c1 = time.perf_counter()
        gip, gind, gdat, _ = coassoc_device(torch.from_numpy(Ls).cuda(), ks)
        eq = bool(np.array_equal(gip.cpu().numpy(), sip) and
                  np.array_equal(gind.cpu().numpy(), sind) and
                  np.array_equal(gdat.cpu().numpy(), scnt))
        out["cpu_baseline"] = {"value": sinc / (c1 - c0) / 1e9, "unit": "Gpair-increments/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{ns} of the {n} points (every partition), C "
                                         "restatement of EAC_CSR's surgical build "
                                         "(sparse.py:154-225, :445-465, :956-1099)",
                               "seconds": c1 - c0, "pair_increments": sinc,
                               "equal_to_gpu": eq}
        log(f"CPU sample {c1 - c0:.2f}s, equal {eq}")
    return out


# Borůvka knob candidates for the batched cut's 669M-entry union graph (a
# few settings: each candidate costs about 0.15 s there)
SL_BATCHED_GRID = {"light_x10": (0, 10, 20), "level_step": (0, 8), "first_batch": (0,),
                   "filter_bpcu": (0, -1)}


def run_sl_batched(args):
    """512 independent symmetric CSR graphs (vertex counts log-uniform in
    [1,000, vmax], 3-D k-NN graphs with k = 0.75 x a target mean degree
    U[8, 64], symmetrized, int32 weights U[1, 256]) as one disjoint union on
    the GPU; one batched lifetime cut (eac.py rule) per step, as an eager
    launch sequence and as one hipGraph replay, against one call per
    graph."""
    import numpy as np
    import torch

    from masters_final_amd import synth
    from masters_final_amd.graph.singlelink import (LifetimeCutGraph, lifetime_cut,
                                                    lifetime_cut_batched)

    G = args.graphs
    log(f"generating {G} k-NN graphs")
    t0 = time.perf_counter()
    b = synth.knn_batch_device(G, (1000, args.vmax), (8, 64), seed=4, wmax=256)
    torch.cuda.synchronize()
    d_dest, d_w, d_ip, goff = b["dest"], b["weight"], b["indptr"], b["goff"]
    N, M = b["n"], int(d_dest.numel())
    log(f"batch: {N} vertices, {M} entries, {time.perf_counter() - t0:.1f}s")
    # Borůvka knobs timed on this union graph (the settings left by another
    # workload's tuning are up to 30% slower here)
    from masters_final_amd.tune import autotune_mst, set_knobs
    if args.knobs:
        knobs = {k: int(v) for k, v in (kv.split("=") for kv in args.knobs.split(","))}
        set_knobs(**knobs)
        tuning = {"knobs": knobs, "ms": None, "cached": None, "source": "--knobs"}
    else:
        tuning = autotune_mst(d_dest, d_w, d_ip, grid=SL_BATCHED_GRID, reps=2,
                              use_cache=not args.retune)
        tuning.pop("candidates", None)
    log(f"tuning: {tuning['knobs']} (cached={tuning['cached']})")
    (lab, gout, info), ms = timed(lambda: lifetime_cut_batched(d_dest, d_w, d_ip, goff, rule=0),
                                  args.steps, args.warmup, torch)
    log(f"batched eager {ms:.2f} ms")
    # dominant kernel of the batched cut from a separate profiled pass (the
    # library's HIP events bracket the Borůvka kernels on their stream)
    from masters_final_amd import _lib as L
    lib = L.load_library()
    L.check(lib.eac_profile_enable(1))
    timed(lambda: lifetime_cut_batched(d_dest, d_w, d_ip, goff, rule=0), args.steps, 0, torch)
    prof = read_prof(lib, L, MST_SLOTS)
    L.check(lib.eac_profile_enable(0))
    dom = max(prof, key=lambda s: prof[s][0])
    tm, nl, nb = prof[dom]
    avg_ms = tm / max(nl, 1)
    bpl = nb / max(nl, 1)
    ach = bpl / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    roof = {"bound": "hbm", "kernel": MST_SLOTS[dom], "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
            "avg_launch_ms": avg_ms, "alg_bytes_per_launch": bpl,
            "kernels_ms_per_step": {MST_SLOTS[s]: prof[s][0] / args.steps for s in prof}}
    # the same cut recorded as one hipGraph (rounds tuned on the first
    # replay); the batch is loaded into the recorded buffers before timing
    cg = LifetimeCutGraph(N + 1, M + 1, G + 1, torch.int32, rule=0, tune=True)
    cg.run(d_dest, d_w, d_ip, goff)
    cg.load(d_dest, d_w, d_ip, goff)
    _, mg = timed(cg.launch, args.steps, args.warmup, torch)
    s_g = cg.summary[:5].cpu().numpy()
    same_graph = (int(s_g[4]) == 0 and bool(torch.equal(cg.labels[:N], lab)) and
                  bool(torch.equal(cg.gout[:G], gout)))
    log(f"graph replay {mg:.2f} ms ({cg.rounds} recorded rounds), equal: {same_graph}")
    # plain: one call per graph
    ip_h = d_ip.cpu().numpy()
    parts = []
    for g in range(G):
        a, z = int(goff[g]), int(goff[g + 1])
        e0, e1 = int(ip_h[a]), int(ip_h[z])
        parts.append((d_dest[e0:e1] - a, d_w[e0:e1], d_ip[a:z + 1] - e0))

    def plain():
        return [lifetime_cut(dd, ww, ii, None, rule=0) for dd, ww, ii in parts]

    rp, mp = timed(plain, 1, 1, torch)
    log(f"per-graph calls {mp:.2f} ms")
    same = True
    lab_h = lab.cpu().numpy()
    gout_h = gout.cpu().numpy()
    for g in range(G):
        a, z = goff[g], goff[g + 1]
        lg = rp[g][0]
        if lg is None:
            same &= bool(gout_h[g, 1] == 1)
        else:
            same &= bool(np.array_equal(lab_h[a:z], lg.cpu().numpy()))
    cpu = None
    if not args.no_cpu:
        # per-graph CPU reference (C restatement of eac.py:199-267 over scipy's
        # Kruskal order) on the first graphs, about 10M entries
        from oracle import oracle as O
        dest_h = d_dest.cpu().numpy()
        w_h = d_w.cpu().numpy()
        done_e, secs, ok, ng = 0, 0.0, True, 0
        for g in range(G):
            a, z = int(goff[g]), int(goff[g + 1])
            e0, e1 = int(ip_h[a]), int(ip_h[z])
            fe = ip_h[a:z] - e0
            od = np.diff(ip_h[a:z + 1])
            c0 = time.perf_counter()
            u, v, w, _ = O.csr_edges(dest_h[e0:e1] - a, w_h[e0:e1], fe, od)
            try:
                nc, want = O.sp_sl_lifetime(u, v, w, z - a)
                ok &= bool(want.size == 0 and gout_h[g, 1] == 1) or \
                    bool(np.array_equal(lab_h[a:z], want))
            except ValueError:
                ok &= bool(gout_h[g, 3] == -1)
            secs += time.perf_counter() - c0
            done_e += e1 - e0
            ng += 1
            if done_e >= 10_000_000:
                break
        cpu = {"value": done_e / secs / 1e6, "unit": "Medges/s", "cores": 1, "kind": "port",
               "sample": f"first {ng} graphs ({done_e} entries), C restatement of the "
                         "per-graph eac.py:199-267 sp_sl_lifetime (Kruskal order)",
               "seconds": secs, "equal_to_gpu": ok}
        log(f"CPU sample {ng} graphs {secs:.2f}s, equal {ok}")
        # all cores: the same per-graph restatement on every graph of the
        # batch, graphs spread over OpenMP threads (largest first); a sample
        # of the graphs checked against the GPU labels
        c0 = time.perf_counter()
        ncs, labs, threads = O.sl_batched_omp(ip_h, dest_h, w_h, goff)
        c1 = time.perf_counter()
        okall = all((ncs[g] == -1 and gout_h[g, 3] == -1) or
                    (gout_h[g, 1] == 1 and ncs[g] == 0) or
                    np.array_equal(lab_h[goff[g]:goff[g + 1]], labs[goff[g]:goff[g + 1]])
                    for g in range(G))
        cpu_all = {"value": M / (c1 - c0) / 1e6, "unit": "Medges/s", "cores": threads,
                   "kind": "port", "sample": f"all {G} graphs ({M} entries), the per-graph C "
                   "restatement on OpenMP threads (oracle/eac_oracle.c orc_sl_batched_omp)",
                   "seconds": c1 - c0, "equal_to_gpu": bool(okall)}
        log(f"all-core CPU {c1 - c0:.2f}s on {threads} threads, equal {okall}")
    else:
        cpu_all = None
    out = base_line(args, M / (mg * 1e-3) / 1e6, "Medges/s", mg, "int32",
                    {"workload": "single_link_batched_knn_graphs", "graphs": G,
                     "v_range": [1000, args.vmax], "deg_range": [8, 64], "w_range": [1, 256],
                     "seed": 4, "n_vertices": N, "m_entries": M,
                     "mean_degree": M / N,
                     "step": "hipGraph replay of the batched lifetime cut (eac.py rule)"})
    out["result"] = {"graph_replay_ms": mg, "batched_eager_ms": ms, "per_graph_calls_ms": mp,
                     "speedup_vs_per_graph": mp / mg, "graph_recorded_rounds": cg.rounds,
                     "graph_equal_to_batched": same_graph, "batched_equal_to_per_graph": same,
                     **info}
    out["roofline"] = roof
    out["tuning"] = tuning
    out["cpu_baseline"] = cpu
    out["cpu_baseline_all_cores"] = cpu_all
    return out


def run_accuracy(args):
    """ConsistencyIndex + HungarianIndex of one ensemble-shaped partition
    (k=2048) against a 32-cluster ground truth over 500,000 samples: the
    contingency table on the GPU, the matchings on the host."""
    import numpy as np
    import torch

    from masters_final_amd import _lib as L
    from masters_final_amd.validation.consistency import ConsistencyIndex, greedy_match_count
    from masters_final_amd.validation.hindex import HungarianIndex, hungarian_match_count

    lib = L.load_library()
    n = args.n if args.n != 1_000_000 else 500_000
    rng = np.random.default_rng(11)
    truth = rng.integers(0, 32, n).astype(np.int32)
    part = ((truth.astype(np.int64) * 64 + rng.integers(0, 64, n)) % 2048).astype(np.int32)
    da, db = torch.from_numpy(part).cuda(), torch.from_numpy(truth).cuda()
    out = torch.empty(2048 * 32 + 1, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()
    sp = L.stream_ptr()

    def kernel():
        L.check(lib.eac_contingency(L.ptr(da), L.ptr(db), n, 0, 2048, 0, 32, L.ptr(out), sp))

    for _ in range(args.warmup):
        kernel()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        kernel()
    e1.record(stream)
    e1.synchronize()
    kms = e0.elapsed_time(e1) / args.steps
    T = out.cpu().numpy()[:-1].reshape(2048, 32)
    want = np.bincount(part.astype(np.int64) * 32 + truth, minlength=2048 * 32).reshape(2048, 32)

    def score():
        return (ConsistencyIndex(N=n).score(da, db), HungarianIndex(n).score(da, db))

    (ci, hi), ms = timed(score, args.steps, args.warmup, torch)

    def cpu_score():  # the same indices with the table counted by numpy
        t = np.bincount(part.astype(np.int64) * 32 + truth, minlength=2048 * 32).reshape(2048, 32)
        return (float(greedy_match_count(t, t.sum(1), t.sum(0))) / n,
                hungarian_match_count(t) / n)

    c0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - c0 < 2.0:
        cpu_res = cpu_score()
        reps += 1
    cpu_s = (time.perf_counter() - c0) / reps
    out_line = base_line(args, n / (ms * 1e-3) / 1e6, "Msamples/s", ms, "int32",
                         {"workload": "accuracy_indices", "n": n, "k1": 2048, "k2": 32,
                          "seed": 11})
    ach = n * 8 / (kms * 1e-3) / 1e9
    out_line["result"] = {"consistency": ci, "hungarian": hi, "table_equal_numpy": bool(
        np.array_equal(T, want)), "equal_to_cpu": cpu_res == (ci, hi), "contingency_ms": kms}
    out_line["roofline"] = {"bound": "hbm", "kernel": "k_contingency", "achieved": ach,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                            "traffic": None, "avg_launch_ms": kms}
    out_line["cpu_baseline"] = {"value": n / cpu_s / 1e6, "unit": "Msamples/s", "cores": 1,
                                "kind": "port", "sample": "whole workload: numpy bincount "
                                "table + the same host matchings"}
    return out_line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mst",
                    choices=["mst", "eac", "kmeans", "coassoc", "sl_batched", "accuracy"])
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--knn", type=int, default=20)
    ap.add_argument("--partitions", type=int, default=256)
    ap.add_argument("--graphs", type=int, default=512)
    ap.add_argument("--vmax", type=int, default=200_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-eac", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload only: skip the other workloads' extra keys")
    ap.add_argument("--knobs", default="",
                    help="fixed tuning knobs for the mst or coassoc workload, e.g. "
                         "light_x10=20,level_step=-1 or co_window=32768,co_unroll=8 (skips the "
                         "autotuner)")
    ap.add_argument("--retune", action="store_true",
                    help="ignore the autotuner cache and time the candidates again")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a 256 MiB elementwise copy kernel 5 times (counter "
                         "calibration: 256 MiB read + 256 MiB written per launch)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            torch.add(x, 1.0, out=y)  # one elementwise kernel per launch
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes": int(x.numel()) * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the library has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)
    fn = {"mst": run_mst, "eac": run_eac, "kmeans": run_kmeans, "coassoc": run_coassoc,
          "sl_batched": run_sl_batched, "accuracy": run_accuracy}[args.workload]
    print(json.dumps(fn(args)), flush=True)


if __name__ == "__main__":
    main()
