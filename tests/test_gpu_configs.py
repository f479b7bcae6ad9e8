"""Parity at the full size of every benchmark workload.

Each test builds the workload exactly as bench.py does (same generators and
seeds), runs the HIP path, and checks it against the CPU oracle where the
oracle finishes in seconds, and through size-independent properties where it
does not:

  eac_small      n=10,000 2-D points, 100 partitions, k in [50, 100]: the whole
                 pipeline (ensemble, co-association, forest, cut) equals the
                 oracle's.
  mst_knn_1m     1,000,000-vertex 3-D k-NN graph (k=20): forest edge set equal
                 to the canonical Kruskal forest, component labels equal the
                 oracle's, forest weight equal to boruvka_minho_seq's.
  kmeans_4m      4,194,304 x 64 points, 2048 centroids: split-BF16 labels equal
                 the FP64 labels on every point and the oracle's on a 20,000
                 point sample.
  coassoc_500k   500,000 x 16 points, 256 partitions, k in [32, 2048]: ensemble
                 partitions of the smallest k equal the oracle's fit; the
                 co-association (7e9 nonzeros, int64 offsets) satisfies the
                 pair-count identity, count bounds, ascending columns above the
                 diagonal, and equals a brute-force count on sampled rows.
  sl_batched     512 k-NN graphs, 1,000 to 200,000 vertices: every sampled
                 graph's labels (the largest and smallest included) equal the
                 per-graph oracle's; the hipGraph replay equals the eager call.
"""
import gc

import numpy as np
import pytest

from masters_final_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    yield t
    gc.collect()
    t.cuda.empty_cache()


def _say(msg):
    print(f"[configs] {msg}", flush=True)


def _free(torch):
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_eac_small_pipeline_equals_oracle(torch):
    from masters_final_amd.EAC.eac import eac_pipeline
    n, P, iters = 10_000, 100, 3
    X, _ = synth.gmm_points(n, 2, 6, 0, mean_scale=10.0, std=1.0)
    rng = np.random.default_rng(0)
    ks = rng.integers(50, 101, P)
    init = np.concatenate([X[rng.choice(n, int(k), replace=False)] for k in ks])
    off = np.concatenate([[0], np.cumsum(ks)])
    _say("eac_small: GPU pipeline n=10000 P=100")
    r = eac_pipeline(torch.from_numpy(X).cuda(), ks, iters=iters, precision="bf16x3",
                     init=torch.from_numpy(init).cuda())
    got_ens = r["ensemble_labels"].cpu().numpy()
    _say("eac_small: oracle")
    labs = np.stack([O.kmeans_fit(X, init[off[p]:off[p + 1]], max_iter=iters,
                                  reseed_mode=1)[0] for p in range(P)])
    assert np.array_equal(got_ens, labs)
    ip, ind, cnt = O.coassoc_condensed(labs, ks)
    assert r["nnz"] == ind.size and r["max_count"] == int(cnt.max())
    w = (cnt.max() + 1 - cnt).astype(np.float64)
    rows = np.repeat(np.arange(n), np.diff(ip))
    nc, lab = O.sp_sl_lifetime(rows, ind, w, n, max_dist=P)
    assert r["n_clusters"] == nc
    got = r["labels"].cpu().numpy() if r["labels"] is not None else np.zeros(0)
    assert np.array_equal(got, lab)


def test_mst_knn_1m_forest_and_labels_equal_oracle(torch):
    from masters_final_amd.graph.mst import boruvka_mst
    _say("mst_knn_1m: generating the 1,000,000-vertex k-NN graph (k=20)")
    g = synth.knn_graph(n=1_000_000, k=20, seed=1)
    n = g["n"]
    fe, od = g["indptr"][:-1], np.diff(g["indptr"])
    _say(f"mst_knn_1m: GPU forest, {g['dest'].size} entries")
    r = boruvka_mst(torch.from_numpy(g["dest"]).cuda(), torch.from_numpy(g["weight"]).cuda(),
                    torch.from_numpy(g["indptr"]).cuda(), None, want_ids=True, want_weights=True,
                    want_labels=True)
    nm = r["n_mst"]
    ids = r["ids"][:nm].cpu().numpy()
    labels = r["labels"].cpu().numpy()
    rows = np.repeat(np.arange(n, dtype=np.int64), od)
    a, b = rows[ids], g["dest"][ids].astype(np.int64)
    assert np.all(a < b)  # ids are taken in the smaller endpoint's row
    got = np.sort(a * n + b)
    _say("mst_knn_1m: oracle (canonical Kruskal, components, boruvka_minho_seq)")
    u, v, w, _ = O.csr_edges(g["dest"], g["weight"], fe, od)
    inf, _ = O.kruskal_canonical(u, v, w, n)
    want = np.sort(u[inf] * n + v[inf])
    assert np.array_equal(got, want)
    nc, lab = O.components(u, v, n)
    assert r["n_components"] == nc and np.array_equal(labels, lab)
    mst, nm_seq, _ = O.boruvka_minho_seq(g["dest"], g["weight"], fe, od)
    wt = int(g["weight"][mst[:nm_seq]].astype(np.int64).sum())
    assert nm_seq == nm and wt == int(w[inf].sum())
    assert int(r["weights"][:nm].to(torch.int64).sum().item()) == wt
    del r
    _free(torch)


def test_kmeans_4m_d64_labels_equal_fp64(torch):
    from masters_final_amd.cluster.KMeans import label_device
    n, d, k = 4_194_304, 64, 2048
    _say("kmeans_4m: generating 4,194,304 x 64 points")
    X, _ = synth.gmm_points(n, d, 2048, 2, mean_std=2.0, std=1.0)
    rng = np.random.default_rng(2)
    Cc = X[rng.choice(n, k, replace=False)].copy()
    dX, dC = torch.from_numpy(X).cuda(), torch.from_numpy(Cc).cuda()
    _say("kmeans_4m: split-BF16 and FP64 labeling")
    lab, dist, nres = label_device(dX, dC, precision="bf16x3")
    lab64, dist64, _ = label_device(dX, dC, precision="fp64")
    assert torch.equal(lab, lab64) and torch.equal(dist, dist64)
    assert 0 < nres < n // 20
    samp = np.sort(np.random.default_rng(12).choice(n, 20_000, replace=False))
    samp = np.unique(np.concatenate([samp, [0, n - 1]]))
    _say(f"kmeans_4m: oracle on {samp.size} points ({nres} re-scored on the GPU)")
    want, wd = O.kmeans_label_f64(X[samp], Cc)
    assert np.array_equal(lab[0].cpu().numpy()[samp], want)
    assert np.array_equal(dist[0].cpu().numpy()[samp], wd.astype(np.float32))
    del dX, dC, lab, dist, lab64, dist64
    _free(torch)


def _brute_row(L, i):
    """Co-association row i above the diagonal by direct comparison of the
    labels of every partition: (columns j > i, counts)."""
    cnt = (L[:, i + 1:] == L[:, i:i + 1]).sum(0)
    j = np.nonzero(cnt)[0]
    return j + i + 1, cnt[j]


def test_coassoc_500k_p256_invariants_and_sampled_rows(torch):
    from masters_final_amd.EAC.sparse import coassoc_device
    from masters_final_amd.utils.partition import generate_ensemble_batched
    n, P = 500_000, 256
    _say("coassoc_500k: generating 500,000 x 16 anisotropic points")
    X, _ = synth.gmm_points(n, 16, 32, 3, mean_scale=20.0, anisotropic=(0.2, 3.0))
    rng = np.random.default_rng(3)
    ks = rng.integers(32, 2049, P)
    dX = torch.from_numpy(X).cuda()
    _say("coassoc_500k: batched ensemble (256 partitions)")
    labels, fit = generate_ensemble_batched(dX, ks, iters=3, seed=3, precision="bf16x3")
    L = labels.cpu().numpy()
    # the partitions of the smallest k against the oracle's fit (same initial
    # centroids: the generator's sample indices)
    rs = np.random.default_rng(3)
    idx = [rs.choice(n, int(k), replace=False) for k in ks]
    for p in np.argsort(ks, kind="stable")[:2]:
        _say(f"coassoc_500k: oracle fit of partition {p} (k={ks[p]})")
        want = O.kmeans_fit(X, X[idx[p]], max_iter=3, reseed_mode=1)[0]
        assert np.array_equal(L[p], want), p
    _say("coassoc_500k: co-association build")
    ip, ind, dat, mx = coassoc_device(labels, ks)
    nnz = int(ind.numel())
    ip_h = ip.cpu().numpy()
    assert ip_h[0] == 0 and ip_h[-1] == nnz and np.all(np.diff(ip_h) >= 0)
    assert nnz > 2**31  # int64 offsets in use
    pairs = 0
    for p in range(P):
        s = np.bincount(L[p], minlength=int(ks[p])).astype(np.int64)
        pairs += int((s * (s - 1) // 2).sum())
    _say(f"coassoc_500k: {nnz} nonzeros, {pairs} pairs; full-size invariants")
    total, lo, hi, bad = 0, P + 1, 0, 0
    step = 1 << 30
    starts = torch.from_numpy(ip_h[1:-1]).cuda()
    for c0 in range(0, nnz, step):
        c1 = min(nnz, c0 + step)
        dch = dat[c0:c1]
        total += int(dch.sum(dtype=torch.int64).item())
        lo = min(lo, int(dch.min().item()))
        hi = max(hi, int(dch.max().item()))
        e1 = min(nnz, c1 + 1)
        ich = ind[c0:e1]
        down = (ich[1:] <= ich[:-1])  # position q: entry c0+q+1 not above entry c0+q
        # a row start may go down; every other step must go up
        rs_in = starts[(starts > c0) & (starts < e1)] - c0 - 1
        down[rs_in] = False
        bad += int(down.sum().item())
        del dch, ich, down
    assert total == pairs
    assert lo >= 1 and hi == mx and hi <= P
    assert bad == 0
    nonempty = np.nonzero(np.diff(ip_h))[0]
    first = ind[torch.from_numpy(ip_h[nonempty]).cuda()].cpu().numpy()
    last = ind[torch.from_numpy(ip_h[nonempty + 1] - 1).cuda()].cpu().numpy()
    assert np.all(first > nonempty) and np.all(last < n)
    sample = np.unique(np.concatenate([[0, n - 2, n - 1], np.random.default_rng(33).choice(
        n, 100, replace=False)]))
    _say(f"coassoc_500k: brute-force count of {sample.size} rows")
    for i in sample:
        a, b = int(ip_h[i]), int(ip_h[i + 1])
        cols, cnt = _brute_row(L, int(i))
        assert np.array_equal(ind[a:b].cpu().numpy(), cols), i
        assert np.array_equal(dat[a:b].cpu().numpy(), cnt), i
    del ip, ind, dat, labels, dX
    _free(torch)


def test_sl_batched_512_knn_graphs_equal_per_graph_oracle(torch):
    from masters_final_amd.graph.singlelink import LifetimeCutGraph, lifetime_cut_batched
    G = 512
    _say("sl_batched: generating 512 k-NN graphs")
    b = synth.knn_batch_device(G, (1000, 200_000), (8, 64), seed=4, wmax=256)
    goff = b["goff"]
    N, M = b["n"], int(b["dest"].numel())
    _say(f"sl_batched: {N} vertices, {M} entries; batched cut")
    lab, gout, info = lifetime_cut_batched(b["dest"], b["weight"], b["indptr"], goff, rule=0)
    cg = LifetimeCutGraph(N + 1, M + 1, G + 1, torch.int32, rule=0, tune=True)
    glab, ggout, _ = cg.run(b["dest"], b["weight"], b["indptr"], goff)
    assert torch.equal(glab[:N], lab) and torch.equal(ggout[:G], gout)
    cg.close()
    lab_h, gout_h = lab.cpu().numpy(), gout.cpu().numpy()
    nv = np.diff(goff)
    pick = set(np.random.default_rng(44).choice(G, 20, replace=False).tolist())
    pick |= {int(np.argmax(nv)), int(np.argmin(nv))}
    ip = b["indptr"]
    for g in sorted(pick):
        a, z = int(goff[g]), int(goff[g + 1])
        e0, e1 = int(ip[a].item()), int(ip[z].item())
        dest = b["dest"][e0:e1].cpu().numpy().astype(np.int64) - a
        w = b["weight"][e0:e1].cpu().numpy()
        ipg = ip[a:z + 1].cpu().numpy() - e0
        u, v, ww, _ = O.csr_edges(dest, w, ipg[:-1], np.diff(ipg))
        try:
            nc, want = O.sp_sl_lifetime(u, v, ww, z - a)
        except ValueError:
            assert gout_h[g, 3] == -1, g
            continue
        if want.size == 0:
            assert gout_h[g, 1] == 1, g
            continue
        assert gout_h[g, 0] == nc, g
        assert np.array_equal(lab_h[a:z], want), g
    _say(f"sl_batched: {len(pick)} graphs equal to the oracle")
    del b, lab, gout, glab, ggout
    _free(torch)
