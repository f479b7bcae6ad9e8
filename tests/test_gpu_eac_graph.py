"""The ensemble + co-association recorded in one HIP graph
(EnsembleCoassocGraph: eac_kmeans_fit_graph + eac_coassoc_graph) gives the
eager pipeline's ensemble, dissimilarity matrix (EAC_CSR + make_diassoc,
sparse.py:154-311) and final clustering (eac.py:199-267) on every replay,
with eager work and synchronisations between replays; an undersized
capacity is reported instead of written past."""
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


def _init(X, ks, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([X[rng.choice(X.shape[0], int(k), replace=False)] for k in ks])


@pytest.mark.parametrize("n,d,P,kmin,kmax,iters,prec", [
    (3000, 2, 30, 10, 40, 3, "bf16x3"),    # the FP64 direct labeler (d < 16)
    (2500, 20, 12, 8, 60, 3, "bf16x3"),    # split-BF16 labeling with re-scored flags
])
def test_graph_pipeline_equals_eager(torch, n, d, P, kmin, kmax, iters, prec):
    from masters_final_amd.EAC.eac import EnsembleCoassocGraph, eac_pipeline, eac_pipeline_graph
    X, _ = synth.gmm_points(n, d, 6, 5, mean_scale=10.0)
    X = X.astype(np.float32)
    ks = list(np.random.default_rng(6).integers(kmin, kmax + 1, P))
    dX = torch.from_numpy(X).cuda()
    g = EnsembleCoassocGraph(dX, ks, iters=iters, precision=prec)
    for seed in (1, 2, 3):
        init = torch.from_numpy(_init(X, ks, seed)).cuda()
        ref = eac_pipeline(dX, ks, iters=iters, precision=prec, init=init)
        torch.cuda.synchronize()
        ip, ind, dat, mx, fit = g.run(init)
        from masters_final_amd.EAC.sparse import coassoc_device
        rip, rind, rdat, rmx = coassoc_device(ref["ensemble_labels"], ks, w_offset="max+1")
        assert torch.equal(fit["labels"], ref["ensemble_labels"])
        assert torch.equal(ip, rip) and torch.equal(ind, rind) and torch.equal(dat, rdat)
        assert mx == rmx == ref["max_count"]
        got = eac_pipeline_graph(g, init)
        assert got["n_clusters"] == ref["n_clusters"]
        assert got["n_rescored"] == ref["n_rescored"]
        if ref["labels"] is None:
            assert got["labels"] is None
        else:
            assert torch.equal(got["labels"], ref["labels"])
        _ = dX.sum().item()  # eager work and a synchronisation between replays


def test_graph_pipeline_reports_small_capacity(torch):
    from masters_final_amd.EAC.eac import EnsembleCoassocGraph
    X, _ = synth.gmm_points(1500, 2, 3, 7, mean_scale=10.0)
    X = X.astype(np.float32)
    ks = [4, 5, 6, 7, 8]
    g = EnsembleCoassocGraph(torch.from_numpy(X).cuda(), ks, iters=3, cap_nnz=1000)
    with pytest.raises(RuntimeError, match="error bits"):
        g.run(torch.from_numpy(_init(X, ks, 1)).cuda())


@pytest.mark.parametrize("n,d,P,kmin,kmax,nclu", [
    (3000, 2, 30, 10, 40, 0),     # lifetime rule picks the number of clusters
    (2500, 20, 12, 8, 60, 0),
    (2000, 2, 16, 5, 30, 4),      # a requested number of clusters
])
def test_recorded_pipeline_equals_eager(torch, n, d, P, kmin, kmax, nclu):
    """EACPipelineGraph (ensemble, co-association and the lifetime cut in
    two recorded graphs, one read-back) gives eac_pipeline's clustering on
    every replay, including replays whose matrix differs from the one the
    capacity and the round count were sized on."""
    from masters_final_amd.EAC.eac import EACPipelineGraph, eac_pipeline
    X, _ = synth.gmm_points(n, d, 6, 5, mean_scale=10.0)
    X = X.astype(np.float32)
    ks = list(np.random.default_rng(6).integers(kmin, kmax + 1, P))
    dX = torch.from_numpy(X).cuda()
    pg = EACPipelineGraph(dX, ks, torch.from_numpy(_init(X, ks, 0)).cuda(), n_clusters=nclu)
    assert 0 < pg.rounds <= pg.bound_rounds
    for seed in (1, 2, 3, 0):
        init = torch.from_numpy(_init(X, ks, seed)).cuda()
        ref = eac_pipeline(dX, ks, init=init, n_clusters=nclu)
        got = pg.run(init)
        for key in ("n_clusters", "n_rescored", "nnz", "max_count"):
            assert got[key] == ref[key], key
        assert torch.equal(got["ensemble_labels"], ref["ensemble_labels"])
        if ref["labels"] is None:
            assert got["labels"] is None
        else:
            assert torch.equal(got["labels"], ref["labels"])
        _ = dX.sum().item()  # eager work and a synchronisation between replays


def test_recorded_pipeline_rounds_overflow_is_rerun(torch):
    """A replay that needs more Borůvka rounds than were recorded is
    detected on the device and cut again at the bound, same result."""
    from masters_final_amd.EAC.eac import EACPipelineGraph, eac_pipeline
    X, _ = synth.gmm_points(2000, 2, 5, 5, mean_scale=10.0)
    X = X.astype(np.float32)
    ks = list(np.random.default_rng(3).integers(5, 40, 20))
    dX = torch.from_numpy(X).cuda()
    init = torch.from_numpy(_init(X, ks, 1)).cuda()
    pg = EACPipelineGraph(dX, ks, init)
    pg._capture(1)  # one round: every real matrix needs more
    got = pg.run(init)
    assert pg.rounds == pg.bound_rounds
    ref = eac_pipeline(dX, ks, init=init)
    assert got["n_clusters"] == ref["n_clusters"]
    if ref["labels"] is not None:
        assert torch.equal(got["labels"], ref["labels"])


def test_recorded_pipeline_reports_small_capacity(torch):
    """A matrix larger than the recorded capacity is reported (the
    co-association's capacity error bits), never cut from a truncated
    matrix."""
    from masters_final_amd.EAC.eac import EACPipelineGraph
    X, _ = synth.gmm_points(1500, 2, 3, 7, mean_scale=10.0)
    X = X.astype(np.float32)
    ks = [4, 5, 6, 7, 8]
    init = torch.from_numpy(_init(X, ks, 1)).cuda()
    with pytest.raises(RuntimeError, match="error bits [23]"):
        EACPipelineGraph(torch.from_numpy(X).cuda(), ks, init, margin=0.5)
