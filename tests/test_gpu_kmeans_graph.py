"""The batched K-Means fit recorded in a HIP graph (eac_kmeans_fit_graph:
no host read-back; flags re-scored from a device count; empty clusters
reseeded on the device) gives fit_batched's labels, dists, centroids,
iteration counts and tallies (KMeans.py:117-189 per partition), on every
replay, with eager work and synchronisations between the replays."""
import numpy as np
import pytest

from masters_final_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def _inits(X, ks, seeds, dup=0):
    rng_inits = []
    for s in seeds:
        rng = np.random.default_rng(s)
        idx = np.concatenate([rng.choice(X.shape[0], int(k), replace=False) for k in ks])
        init = X[idx].copy()
        if dup:  # repeated centroids: the later copies start empty
            off = 0
            for k in ks:
                init[off + 1:off + 1 + min(dup, int(k) - 1)] = init[off]
                off += int(k)
        rng_inits.append(init)
    return rng_inits


def _check(torch, X, ks, max_iter, precision, seeds, dup=0, expect_reseed=False):
    from masters_final_amd.cluster.KMeans import BatchedFitGraph, fit_batched
    dX = torch.from_numpy(X).cuda()
    gfit = BatchedFitGraph(dX, ks, max_iter=max_iter, tol=1e-4, precision=precision)
    reseeded = 0
    for init in _inits(X, ks, seeds, dup):
        ref = fit_batched(dX, torch.from_numpy(init).cuda(), ks, max_iter=max_iter, tol=1e-4,
                          precision=precision)
        torch.cuda.synchronize()
        got = gfit.run(torch.from_numpy(init).cuda()).result()
        assert torch.equal(got["labels"], ref["labels"])
        assert torch.equal(got["dists"], ref["dists"])
        assert torch.equal(got["centroids"], ref["centroids"])
        assert np.array_equal(got["iters"], ref["iters"])
        assert got["n_rescored"] == ref["n_rescored"]
        assert got["n_reseeded"] == ref["n_reseeded"]
        reseeded += got["n_reseeded"]
        _ = dX.sum().item()  # eager work and a synchronisation between replays
    if expect_reseed:
        assert reseeded > 0


def test_graph_fit_direct_path_2d(torch):
    X, _ = synth.gmm_points(6000, 2, 8, 0, mean_scale=10.0)
    ks = list(np.random.default_rng(1).integers(5, 40, 24))
    _check(torch, X.astype(np.float32), ks, 4, "bf16x3", seeds=[3, 4, 5])


def test_graph_fit_split_bf16_flags_and_sorted_rescoring(torch):
    """d = 40: the split-BF16 labeling flags points; the flags are counted,
    sorted by partition and re-scored from device counts."""
    X, _ = synth.gmm_points(5000, 40, 12, 1, mean_scale=3.0)
    ks = list(np.random.default_rng(2).integers(20, 120, 9))
    _check(torch, X.astype(np.float32), ks, 3, "bf16x3", seeds=[7, 8])


def test_graph_fit_reseeds_empty_clusters_on_device(torch):
    X, _ = synth.gmm_points(3000, 3, 6, 2, mean_scale=8.0)
    ks = list(np.random.default_rng(3).integers(6, 30, 12))
    _check(torch, X.astype(np.float32), ks, 5, "fp64", seeds=[11, 12, 13], dup=4,
           expect_reseed=True)


def test_graph_fit_stops_partitions_early(torch):
    """tol = 1e-4 with max_iter 12: partitions stop at different
    iterations; the graph still runs 12 labelings and matches."""
    X, _ = synth.gmm_points(4000, 2, 4, 3, mean_scale=30.0)
    ks = [2, 3, 4, 5, 8, 16, 30]
    _check(torch, X.astype(np.float32), ks, 12, "fp32", seeds=[21])
