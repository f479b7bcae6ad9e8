"""Dense small-n EAC path on the GPU (EAC_FULL, SL, clusters_from_SL,
full_sl_lifetime, EAC(sparse=False)) against the oracle (orc_coassoc_full,
orc_slink with the reference's argsort1D, orc_labels_from_Z,
orc_full_sl_lifetime; pinned in tests/test_oracle_dense.py) and the
reference's own single-linkage known answers
(tests/golden/hierarchy_single.json)."""
import json
from pathlib import Path

import numpy as np
import pytest
from scipy.spatial.distance import pdist

from masters_final_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "hierarchy_single.json").read_text())


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def _labels(rng, n, P, kmin, kmax):
    return np.stack([rng.integers(0, rng.integers(kmin, kmax + 1), n)
                     for _ in range(P)]).astype(np.int32)


def _same_partition(a, b):
    _, ia = np.unique(np.asarray(a), return_inverse=True)
    _, ib = np.unique(np.asarray(b), return_inverse=True)
    pairs = set(zip(ia.tolist(), ib.tolist()))
    return len(pairs) == len(set(ia.tolist())) == len(set(ib.tolist()))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int32, np.float32])
@pytest.mark.parametrize("condensed", [True, False])
def test_coassoc_dense_equals_oracle(torch, dtype, condensed):
    from masters_final_amd.EAC.full import coassoc_dense_device
    rng = np.random.default_rng(1)
    n, P = 333, 37
    L = _labels(rng, n, P, 2, 9)
    L[4, 10:40] = -1  # samples not listed in one partition
    want = O.coassoc_full(L, condensed=condensed)
    got, mx = coassoc_dense_device(L, condensed=condensed, dtype=dtype)
    assert np.array_equal(got.cpu().numpy().astype(np.int64), want)
    assert mx == want.max()


def test_coassoc_dense_uint8_wraps(torch):
    """300 identical partitions: every count is 300, stored as 300 mod 256
    in uint8 (the reference's uint8 += 1 wraps)."""
    from masters_final_amd.EAC.full import EAC_FULL
    n, P = 70, 300
    L = np.tile((np.arange(n) % 3)[None], (P, 1)).astype(np.int32)
    m = EAC_FULL(n, dtype=np.uint8, condensed=True)
    m.update_labels(L)
    want = (O.coassoc_full(L, condensed=True) % 256).astype(np.uint8)
    assert np.array_equal(m.coassoc.cpu().numpy(), want)
    m.make_diassoc()  # max stored value 44 (300 mod 256) -> 45 - x, wrapping
    assert np.array_equal(m.coassoc.cpu().numpy(), (45 - want.astype(np.int64)) % 256)


def test_eac_full_degree_and_diassoc(torch):
    from masters_final_amd.EAC.full import EAC_FULL
    from masters_final_amd.utils.partition import convertClusterStringToIndex
    rng = np.random.default_rng(2)
    n, P = 500, 12
    L = _labels(rng, n, P, 4, 30)
    ens = [convertClusterStringToIndex(lab) for lab in L]
    relab = np.stack([np.searchsorted(np.unique(lab), lab) for lab in L])
    for condensed in (True, False):
        m = EAC_FULL(n, dtype=np.int32, condensed=condensed)
        m.update_ensemble(ens)
        want = O.coassoc_full(relab, condensed=True)
        sq = np.zeros((n, n), np.int64)
        sq[np.triu_indices(n, 1)] = want
        sq = sq + sq.T
        assert np.array_equal(m.todense(), sq)
        m.get_degree()
        if condensed:
            deg = np.array([(sq[i, i + 1:] != 0).sum() for i in range(n)])
        else:
            deg = (sq != 0).sum(1)
        assert np.array_equal(m.degree, deg) and m.nnz == int(deg.sum())
        m.make_diassoc()
        assert np.array_equal(m.coassoc.cpu().numpy(),
                              (want.max() + 1 - (want if condensed else sq)).astype(np.int32))


def test_slink_ytdist_golden(torch):
    from masters_final_amd.cluster.linkage import SL
    Z = SL(np.array(GOLD["ytdist"], np.float32), 6)
    assert Z.dtype == np.float32
    assert np.array_equal(Z, np.array(GOLD["linkage_ytdist_single"], np.float32))


@pytest.mark.parametrize("t", ["8.0", "4.0", "1.0"])
def test_clusters_from_sl_q_x_golden(torch, t):
    from masters_final_amd.cluster.linkage import SL, clusters_from_SL
    X = np.array(GOLD["Q_X"])
    Z = SL(pdist(X).astype(np.float32), X.shape[0])
    track = np.zeros(X.shape[0], np.int32)
    clusters_from_SL(Z, track, int(float(t)))
    assert _same_partition(track, GOLD["fcluster_maxclust"][t])
    assert np.array_equal(track, O.labels_from_Z(Z, int(float(t))))


@pytest.mark.parametrize("seed,n", [(3, 64), (4, 1000), (5, 2500)])
def test_slink_equals_oracle_distinct_heights(torch, seed, n):
    """Distinct float32 heights: the single-linkage Z is unique, so the GPU's
    equals the oracle's SLINK + argsort1D Z entry for entry; labels of every
    tenth cut equal the oracle's."""
    from masters_final_amd.cluster.linkage import SL, labels_from_Z_device
    X = np.random.default_rng(seed).random((n, 4))
    d = pdist(X).astype(np.float32)
    Zo = O.slink(d.astype(np.float64), n)
    Z = SL(d, n)
    assert np.array_equal(Z, Zo.astype(np.float32))
    for nc in range(1, n + 1, max(1, n // 10)):
        got = labels_from_Z_device(Z, nc).cpu().numpy()
        assert np.array_equal(got, O.labels_from_Z(Zo, nc)), nc


@pytest.mark.parametrize("seed,n,P", [(6, 400, 20), (7, 2000, 50)])
def test_full_sl_lifetime_equals_oracle_on_eac(torch, seed, n, P):
    """uint8 EAC dissimilarities (many equal heights): heights, cluster count
    and labels of the lifetime cut equal the oracle's (the cut takes every
    merge at or below a height, which no tie order changes)."""
    from masters_final_amd.EAC.full import EAC_FULL, full_sl_lifetime_device
    X, _ = synth.gmm_points(n, 2, 5, seed, mean_scale=8.0)
    rng = np.random.default_rng(seed)
    from masters_final_amd.cluster.KMeans import label_device
    L = np.stack([label_device(X, X[rng.choice(n, int(k), replace=False)], precision="fp64")[0][0]
                  .cpu().numpy() for k in rng.integers(5, 30, P)]).astype(np.int32)
    m = EAC_FULL(n, dtype=np.uint8, condensed=True)
    m.update_labels(L)
    m.make_diassoc()
    dis = m.coassoc.cpu().numpy()
    try:
        nc_o, lab_o, Zo = O.full_sl_lifetime(dis.astype(np.float64), n)
    except ValueError:
        with pytest.raises(ValueError):
            full_sl_lifetime_device(m.coassoc, n)
        return
    nc, lab, Z = full_sl_lifetime_device(m.coassoc, n, want_Z=True)
    assert nc == nc_o
    assert np.array_equal(Z[:, 2].cpu().numpy(), Zo[:, 2].astype(np.float32))
    assert Z[-1, 3].item() == n  # the last merge holds every sample
    if lab_o.size == 0:
        assert lab is None
    else:
        assert np.array_equal(lab.cpu().numpy(), lab_o)


def test_eac_driver_dense_equals_sparse_when_connected(torch):
    """EAC(sparse=False) and EAC(sparse=True) on the same ensemble: on a
    connected co-association graph both cuts give the same labels."""
    from masters_final_amd.EAC.eac import EAC
    from masters_final_amd.utils.partition import convertClusterStringToIndex
    rng = np.random.default_rng(9)
    n, P = 600, 25
    L = _labels(rng, n, P, 2, 6)
    ens = [convertClusterStringToIndex(lab) for lab in L]
    relab = np.stack([np.searchsorted(np.unique(lab), lab) for lab in L])
    ip, ind, _ = O.coassoc_condensed(relab, relab.max(1) + 1)
    rows = np.repeat(np.arange(n), np.diff(ip))
    assert O.components(rows, ind, n)[0] == 1
    dense = EAC(n, sparse=False, n_partitions=P)
    dense.buildMatrix(ens)
    a = dense.finalClustering()
    sparse = EAC(n, sparse=True, n_partitions=P)
    sparse.buildMatrix(ens)
    b = sparse.finalClustering()
    assert dense.n_fclusts == sparse.n_fclusts and np.array_equal(a, b)
    cond = O.coassoc_full(relab, condensed=True)
    nc, lab, _ = O.full_sl_lifetime((cond.max() + 1 - cond).astype(np.float64), n)
    assert nc == dense.n_fclusts and np.array_equal(a, lab)


def test_full_sl_lifetime_errors_and_one_cluster(torch):
    from masters_final_amd.EAC.full import full_sl_lifetime
    with pytest.raises(ValueError):  # all heights distinct: numpy raises
        full_sl_lifetime(np.array(GOLD["ytdist"], np.float32), 6)
    nc, lab = full_sl_lifetime(np.ones(10, np.float32), 5)
    assert nc == 1 and lab.size == 0
    nc, lab = full_sl_lifetime(np.array(GOLD["ytdist"], np.float32), 6, n_clusters=2)
    Zo = O.slink(np.array(GOLD["ytdist"]), 6)
    assert nc == 2 and np.array_equal(lab, O.labels_from_Z(Zo, 2))
