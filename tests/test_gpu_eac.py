"""GPU co-association builder and EAC end-to-end against the oracle."""
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
    return t


def _dense(labels):
    P, n = labels.shape
    C = np.zeros((n, n), np.int64)
    for p in range(P):
        for c in np.unique(labels[p]):
            if c < 0:
                continue
            idx = np.nonzero(labels[p] == c)[0]
            np.add.at(C, (idx[:, None], idx[None, :]), 1)
    np.fill_diagonal(C, 0)
    return C


@pytest.mark.parametrize("n,P,kmin,kmax,seed", [(500, 10, 3, 20, 0), (3000, 40, 20, 60, 1),
                                                (20000, 8, 2, 6, 2)])
def test_coassoc_condensed_vs_oracle(torch, n, P, kmin, kmax, seed):
    from masters_final_amd.EAC.sparse import coassoc_device
    rng = np.random.default_rng(seed)
    ks = rng.integers(kmin, kmax + 1, P)
    labels = np.stack([rng.integers(0, k, n) for k in ks]).astype(np.int32)
    indptr, ind, cnt = O.coassoc_condensed(labels, ks) if n <= 3000 else (None, None, None)
    gi, gc, gd, mx = coassoc_device(labels, ks)
    gi, gc, gd = gi.cpu().numpy(), gc.cpu().numpy(), gd.cpu().numpy()
    if indptr is not None:
        assert np.array_equal(gi, indptr) and np.array_equal(gc, ind) and np.array_equal(gd, cnt)
    if n <= 3000:
        D = np.triu(_dense(labels), 1)
        assert mx == D.max()
    # rows sorted, columns above the diagonal, counts within [1, P]
    rows = np.repeat(np.arange(n), np.diff(gi))
    assert np.all(gc > rows) and np.all((gd >= 1) & (gd <= P))
    assert np.all(np.diff(gc)[np.diff(rows) == 0] > 0)


@pytest.mark.parametrize("window", [0, 4096])
def test_coassoc_unroll_knob_same_matrix(torch, window):
    """The tuning knobs (LDS window, member loads in flight, row ranges)
    change speed only: every setting gives the oracle's matrix."""
    from masters_final_amd.EAC.sparse import coassoc_device
    from masters_final_amd.tune import get_knobs, set_knobs
    rng = np.random.default_rng(17)
    n, P = 9000, 24
    ks = rng.integers(3, 40, P)
    labels = np.stack([rng.integers(0, k, n) for k in ks]).astype(np.int32)
    saved = get_knobs()
    try:
        outs = []
        for u, ranges in ((4, 0), (8, 1), (4, 3)):
            set_knobs(co_window=window, co_unroll=u, co_ranges=ranges)
            gi, gc, gd, mx = coassoc_device(labels, ks, full_rows=(window == 0))
            outs.append((gi.cpu().numpy(), gc.cpu().numpy(), gd.cpu().numpy(), mx))
    finally:
        set_knobs(**saved)
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)
    ip, ind, cnt = O.coassoc_condensed(labels, ks) if window else (None, None, None)
    if ip is not None:
        assert np.array_equal(outs[1][0], ip) and np.array_equal(outs[1][1], ind)
        assert np.array_equal(outs[1][2], cnt)


def test_coassoc_full_rows_and_dissim(torch):
    from masters_final_amd.EAC.sparse import coassoc_device
    rng = np.random.default_rng(5)
    n, P = 1500, 12
    ks = rng.integers(2, 9, P)
    labels = np.stack([rng.integers(0, k, n) for k in ks]).astype(np.int32)
    labels[3, :100] = -1  # samples left out of one partition
    D = _dense(labels)
    gi, gc, gd, mx = coassoc_device(labels, ks, full_rows=True, w_offset="max+1")
    gi, gc, gd = gi.cpu().numpy(), gc.cpu().numpy(), gd.cpu().numpy()
    r, c = np.nonzero(D)
    assert np.array_equal(gc, c) and np.array_equal(gi[1:], np.cumsum(np.bincount(r, minlength=n)))
    assert mx == D.max() and np.array_equal(gd, mx + 1 - D[r, c])


def test_eac_csr_class(torch):
    from masters_final_amd.EAC.sparse import EAC_CSR
    from masters_final_amd.utils.partition import convertClusterStringToIndex
    rng = np.random.default_rng(6)
    n, P = 800, 15
    ks = rng.integers(3, 12, P)
    labels = np.stack([rng.integers(0, k, n) for k in ks])
    ens = [convertClusterStringToIndex(l) for l in labels]
    # golden vector of the reference's own test (test_partition.py:82-97)
    out = convertClusterStringToIndex(np.array([1, 1, 1, 3, 2, 4, 2, 4, 3, 3], np.uint32))
    assert [a.tolist() for a in out] == [[0, 1, 2], [4, 6], [3, 8, 9], [5, 7]]
    m = EAC_CSR(n, max_assocs=100, condensed=True, max_assocs_type="linear",
                sort_mode="surgical")
    m.update_ensemble(ens)
    m._condense()
    relab = np.stack([np.searchsorted(np.unique(l), l) for l in labels])
    ip, ind, cnt = O.coassoc_condensed(relab, [len(e) for e in ens])
    csr = m.csr
    assert np.array_equal(csr.indptr, ip) and np.array_equal(csr.indices, ind)
    assert np.array_equal(csr.data, cnt)
    m.make_diassoc()
    assert np.array_equal(m.csr.data, cnt.max() + 1 - cnt)


def _oracle_eac(X, init, ks, iters=3):
    n = X.shape[0]
    off = np.concatenate([[0], np.cumsum(ks)])
    labs = []
    for p, k in enumerate(ks):
        lab, _, _, _ = O.kmeans_fit(X, init[off[p]:off[p + 1]], max_iter=iters, reseed_mode=1)
        labs.append(lab)
    labs = np.stack(labs)
    ip, ind, cnt = O.coassoc_condensed(labs, ks)
    w = (cnt.max() + 1 - cnt).astype(np.float64)
    rows = np.repeat(np.arange(n), np.diff(ip))
    nc, lab = O.sp_sl_lifetime(rows, ind, w, n, max_dist=len(ks))
    return labs, nc, lab


def test_eac_pipeline_matches_oracle(torch):
    from masters_final_amd.EAC.eac import eac_pipeline
    X, _ = synth.gmm_points(2000, 2, 6, 0, mean_scale=10.0)
    rng = np.random.default_rng(0)
    ks = rng.integers(10, 21, 20)
    init = np.concatenate([X[rng.choice(2000, k, replace=False)] for k in ks])
    labs, nc, lab = _oracle_eac(X, init, ks)
    for prec in ("bf16x3", "fp32"):
        r = eac_pipeline(X, ks, iters=3, precision=prec, init=init)
        if prec == "bf16x3":
            assert np.array_equal(r["ensemble_labels"].cpu().numpy(), labs)
            assert r["n_clusters"] == nc
            got = r["labels"].cpu().numpy() if r["labels"] is not None else np.zeros(0)
            assert np.array_equal(got, lab)


def test_seeded_ensemble_init_on_device(torch):
    """generate_ensemble_batched without init draws k distinct rows per
    partition from PCG64(seed) and gathers them where X lives: device and
    host inputs give the same init, and the labels equal the oracle's fits."""
    from masters_final_amd.utils.partition import generate_ensemble_batched
    X, _ = synth.gmm_points(3000, 3, 5, 4, mean_scale=10.0)
    ks = np.array([7, 12, 3, 20])
    rng = np.random.default_rng(11)
    init = np.concatenate([X[rng.choice(3000, int(k), replace=False)] for k in ks])
    off = np.concatenate([[0], np.cumsum(ks)])
    want = np.stack([O.kmeans_fit(X, init[off[p]:off[p + 1]], max_iter=3, reseed_mode=1)[0]
                     for p in range(len(ks))])
    for data in (torch.from_numpy(X).cuda(), X):
        labels, _ = generate_ensemble_batched(data, ks, iters=3, seed=11)
        assert np.array_equal(labels.cpu().numpy(), want)


def test_eac_class_end_to_end(torch):
    from masters_final_amd.EAC.eac import EAC
    from masters_final_amd.utils.partition import convertClusterStringToIndex
    X, _ = synth.gmm_points(1500, 2, 5, 3, mean_scale=10.0)
    rng = np.random.default_rng(3)
    ks = rng.integers(8, 16, 12)
    init = np.concatenate([X[rng.choice(1500, k, replace=False)] for k in ks])
    labs, nc, lab = _oracle_eac(X, init, ks)
    ens = [convertClusterStringToIndex(l) for l in labs]
    e = EAC(1500, sparse=True, n_partitions=len(ks))
    e.buildMatrix(ens)
    got = e.finalClustering()
    assert e.n_fclusts == nc and np.array_equal(got, lab)


@pytest.mark.parametrize("window", [64, 1000, 4096])
def test_coassoc_window_knob_keeps_the_matrix(torch, window):
    """Any co-association window (several windows per row included) gives
    the oracle's matrix."""
    from masters_final_amd import tune
    from masters_final_amd.EAC.sparse import coassoc_device
    rng = np.random.default_rng(window)
    n, P = 3000, 12
    ks = rng.integers(3, 40, P)
    lab = np.stack([rng.integers(0, k, n) for k in ks]).astype(np.int32)
    before = tune.get_knobs()
    try:
        tune.set_knobs(co_window=window)
        ip, ind, dat, mx = coassoc_device(torch.from_numpy(lab).cuda(), ks, w_offset=0)
        wip, wind, wcnt = O.coassoc_condensed(lab, ks)
        assert np.array_equal(ip.cpu().numpy(), wip)
        assert np.array_equal(ind.cpu().numpy(), wind)
        assert np.array_equal(dat.cpu().numpy(), wcnt)
    finally:
        tune.set_knobs(**before)
