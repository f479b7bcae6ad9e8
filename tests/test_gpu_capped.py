"""Capacity-capped co-association (the reference's lossy EAC_CSR with
max_assocs, MyML/MyML/EAC/sparse.py:154-225, :368-465, :866-1099) on the GPU
against the oracle's restatement of the surgical build (orc_coassoc_capped,
pinned in tests/test_oracle_capped.py)."""
import numpy as np
import pytest

from masters_final_amd import synth
from masters_final_amd.EAC.sparse import indptr_constant, indptr_linear
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def _labels(rng, n, P, kmin, kmax):
    return np.stack([rng.integers(0, rng.integers(kmin, kmax + 1), n)
                     for _ in range(P)]).astype(np.int32)


def _check(got, want):
    ip, ind, dat, mx, disc = got
    wip, wind, wcnt, wdisc = want
    assert np.array_equal(ip.cpu().numpy(), wip)
    assert np.array_equal(ind.cpu().numpy(), wind)
    assert np.array_equal(dat.cpu().numpy(), wcnt)
    assert disc == wdisc
    assert mx == (int(wcnt.max()) if wcnt.size else 0)


@pytest.mark.parametrize("seed,n,P,kmin,kmax,fac,profile", [
    (0, 2000, 30, 5, 30, 3.0, "linear"),     # the driver's default: 3 x biggest cluster
    (1, 2000, 30, 5, 30, 1.5, "linear"),     # discards
    (2, 3000, 20, 10, 40, 1.0, "constant"),  # heavy discards
    (5, 3000, 20, 10, 40, 0.6, "constant"),  # the first partition overflows a row
    (3, 700, 64, 2, 9, 2.0, "linear"),
])
def test_capped_condensed_equals_oracle(torch, seed, n, P, kmin, kmax, fac, profile):
    from masters_final_amd.EAC.sparse import coassoc_capped_device
    rng = np.random.default_rng(seed)
    L = _labels(rng, n, P, kmin, kmax)
    ks = L.max(1) + 1
    ma = int(fac * max(np.bincount(lab).max() for lab in L))
    off = indptr_linear(n, ma, 0.05, 1.0, 1.0, 0.05) if profile == "linear" \
        else indptr_constant(n, ma)
    try:
        want = O.coassoc_capped(L, off)
    except ValueError:
        with pytest.raises(ValueError):
            coassoc_capped_device(L, ks, off)
        return
    _check(coassoc_capped_device(L, ks, off), want)


def test_capped_full_rows_equals_oracle(torch):
    from masters_final_amd.EAC.sparse import coassoc_capped_device
    rng = np.random.default_rng(4)
    n, P = 1500, 25
    L = _labels(rng, n, P, 4, 20)
    L[7] = rng.integers(0, 3, n)  # the first partition: 3 clusters
    ks = L.max(1) + 1
    ma = int(1.2 * np.bincount(L[7]).max())
    off = indptr_constant(n, ma)
    want = O.coassoc_capped(L, off, full_rows=True, max_assocs=ma)
    assert want[3] > 0
    _check(coassoc_capped_device(L, ks, off, full_rows=True, max_assocs=ma), want)


def test_capped_first_partition_overflow_raises(torch):
    from masters_final_amd.EAC.sparse import coassoc_capped_device
    L = np.zeros((2, 50), np.int32)
    L[1] = np.arange(50) % 5
    with pytest.raises(ValueError):
        coassoc_capped_device(L, [1, 5], indptr_constant(50, 10))


def test_eac_csr_compat_mode_on_gmm_ensemble(torch):
    """EAC_CSR(exact=False) fed with an ensemble in the reference's list
    format (partition.py:73-86) equals the oracle, and reports the
    reference's nnz_discarded; the exact mode of the same matrix equals the
    uncapped oracle."""
    from masters_final_amd.EAC.sparse import EAC_CSR
    from masters_final_amd.cluster.KMeans import label_device
    from masters_final_amd.utils.partition import convertClusterStringToIndex
    X, _ = synth.gmm_points(4000, 2, 6, 21, mean_scale=10.0)
    rng = np.random.default_rng(21)
    P = 24
    ens, labs = [], []
    for p in range(P):
        k = int(rng.integers(5, 40))
        lab = label_device(X, X[rng.choice(X.shape[0], k, replace=False)],
                           precision="fp64")[0][0].cpu().numpy()
        labs.append(np.searchsorted(np.unique(lab), lab))
        ens.append(convertClusterStringToIndex(lab))
    L = np.stack(labs)
    biggest = max(max(a.size for a in part) for part in ens)
    # max_assocs = the biggest cluster (the driver uses 3x): some rows fill up
    m = EAC_CSR(X.shape[0], max_assocs=int(biggest), max_assocs_type="linear",
                condensed=True, sort_mode="surgical", exact=False)
    m.update_ensemble(ens)
    m._condense()
    wip, wind, wcnt, wdisc = O.coassoc_capped(L, m.capacity_offsets())
    csr = m.csr
    assert np.array_equal(csr.indptr, wip) and np.array_equal(csr.indices, wind)
    assert np.array_equal(csr.data, wcnt) and m.nnz_discarded == wdisc > 0
    m.make_diassoc()
    assert np.array_equal(m.csr.data, wcnt.max() + 1 - wcnt)


def test_eac_driver_compat_labels(torch):
    """EAC(sparse=True, sparse_exact=False): the reference's default
    max_assocs (3 x the biggest cluster, linear profile) and its lifetime cut
    on the capped matrix."""
    from masters_final_amd.EAC.eac import EAC
    from masters_final_amd.utils.partition import convertClusterStringToIndex
    rng = np.random.default_rng(8)
    n, P = 3000, 30
    L = _labels(rng, n, P, 8, 30)
    L = np.stack([np.searchsorted(np.unique(lab), lab) for lab in L]).astype(np.int32)
    ens = [convertClusterStringToIndex(lab) for lab in L]
    e = EAC(n, sparse=True, condensed=True, n_partitions=P, sparse_exact=False)
    e.buildMatrix(ens)
    labels = e.finalClustering()
    off = e.coassoc.capacity_offsets()
    ip, ind, cnt, _ = O.coassoc_capped(L, off)
    w = (cnt.max() + 1 - cnt).astype(np.float64)
    rows = np.repeat(np.arange(n), np.diff(ip))
    nc, want = O.sp_sl_lifetime(rows, ind, w, n, max_dist=P)
    assert e.n_fclusts == nc
    assert np.array_equal(labels, want)
