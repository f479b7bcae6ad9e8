"""Pins the oracle's dense EAC path: EAC_FULL counts, argsort1D, SLINK,
labels_from_Z and full_sl_lifetime (MyML/MyML/EAC/full.py:8-181,
scipy_numba/scipy_numba/cluster/_argsort.py, _hierarchy_eac.py:329-927,
MyML/MyML/cluster/linkage.py:303-357, MyML/MyML/EAC/eac.py:337-391).

Checks: the reference's own known answers (ytdist -> linkage_ytdist_single,
Q_X -> fcluster_maxclust, tests/golden/hierarchy_single.json), SciPy's
single linkage (a library routine) on random data, the reference's own
check of argsort1D against np.argsort (_argsort.py:67-74), brute-force
counts, and the identity between the dense and sparse cuts on connected
co-association graphs.
"""
import json
from pathlib import Path

import numpy as np
import pytest
from scipy.cluster.hierarchy import fcluster, linkage
from scipy.spatial.distance import pdist

from oracle import oracle as O

GOLD = json.loads((Path(__file__).parent / "golden" / "hierarchy_single.json").read_text())


def _same_partition(a, b):
    a, b = np.asarray(a), np.asarray(b)
    _, ia = np.unique(a, return_inverse=True)
    _, ib = np.unique(b, return_inverse=True)
    pairs = set(zip(ia.tolist(), ib.tolist()))
    return len(pairs) == len(set(ia.tolist())) == len(set(ib.tolist()))


def test_slink_ytdist_golden():
    Z = O.slink(np.array(GOLD["ytdist"]), 6)
    assert np.array_equal(Z, np.array(GOLD["linkage_ytdist_single"]))


@pytest.mark.parametrize("t", ["8.0", "4.0", "1.0"])
def test_labels_from_Z_q_x_golden(t):
    X = np.array(GOLD["Q_X"])
    Z = O.slink(pdist(X), X.shape[0])
    lab = O.labels_from_Z(Z, int(float(t)))
    assert _same_partition(lab, GOLD["fcluster_maxclust"][t])
    # first-occurrence numbering
    first = [lab[0]]
    for v in lab:
        if v not in first:
            first.append(v)
    assert first == list(range(len(first)))


def test_argsort1d_equals_numpy_on_distinct_values():
    x = np.random.default_rng(0).random(100_000)
    assert np.array_equal(O.argsort1d(x), np.argsort(x))


def test_argsort1d_tie_order_hand_trace():
    # partition of [2, 1, 2, 1] with pivot 2 (index 0): i stops past every
    # value <= 2 (i = 4), j stays at 3, the pivot swaps with index 3; left
    # part [1, 1, 2] (indices 3, 1, 2) with pivot 1 (index 3): i stops at
    # index 2 (value 2), j walks down to 1, pivot swaps with index 1
    assert O.argsort1d(np.array([2.0, 1.0, 2.0, 1.0])).tolist() == [1, 3, 2, 0]


@pytest.mark.parametrize("seed,n", [(0, 40), (1, 200), (2, 700)])
def test_slink_equals_scipy_single_distinct(seed, n):
    X = np.random.default_rng(seed).random((n, 3))
    d = pdist(X)
    Z = O.slink(d, n)
    S = linkage(d, method="single")
    assert np.array_equal(Z[:, 2], S[:, 2])
    assert np.array_equal(Z[:, 3], S[:, 3])
    assert np.array_equal(np.sort(Z[:, :2], axis=1), np.sort(S[:, :2], axis=1))


@pytest.mark.parametrize("seed,n,wmax", [(3, 120, 6), (4, 300, 20)])
def test_slink_ties_heights_and_level_cuts(seed, n, wmax):
    """Integer weights with many ties: the height multiset is SciPy's, and
    cutting after every merge at or below a level gives SciPy's partition at
    that level."""
    d = np.random.default_rng(seed).integers(1, wmax + 1, n * (n - 1) // 2).astype(np.float64)
    Z = O.slink(d, n)
    S = linkage(d, method="single")
    assert np.array_equal(np.sort(Z[:, 2]), S[:, 2])
    for h in np.unique(Z[:, 2]):
        nc = n - int((Z[:, 2] <= h).sum())
        lab = O.labels_from_Z(Z, nc)
        assert _same_partition(lab, fcluster(S, h, criterion="distance"))


def test_coassoc_full_brute_force():
    rng = np.random.default_rng(5)
    n, P = 60, 7
    L = np.stack([rng.integers(0, rng.integers(2, 6), n) for _ in range(P)])
    L[2, :5] = -1
    D = np.zeros((n, n), np.int64)
    for p in range(P):
        for i in range(n):
            for j in range(n):
                if i != j and L[p, i] >= 0 and L[p, i] == L[p, j]:
                    D[i, j] += 1
    assert np.array_equal(O.coassoc_full(L, condensed=False), D)
    iu = np.triu_indices(n, 1)
    assert np.array_equal(O.coassoc_full(L, condensed=True), D[iu])


def test_full_sl_lifetime_equals_sparse_cut_when_connected():
    """On a connected co-association graph the dense matrix adds only
    weight max + 1 pairs, which no minimum spanning tree uses: both cuts see
    the same heights and give the same labels."""
    rng = np.random.default_rng(6)
    n, P = 150, 20
    L = np.stack([rng.integers(0, rng.integers(2, 5), n) for _ in range(P)])
    cond = O.coassoc_full(L, condensed=True)
    mx = cond.max()
    nc, lab, Z = O.full_sl_lifetime((mx + 1 - cond).astype(np.float64), n)
    ip, ind, cnt = O.coassoc_condensed(L, L.max(1) + 1)
    rows = np.repeat(np.arange(n), np.diff(ip))
    ncs, labs = O.sp_sl_lifetime(rows, ind, (cnt.max() + 1 - cnt).astype(np.float64), n,
                                 max_dist=P)
    assert O.components(rows, ind, n)[0] == 1
    assert nc == ncs and np.array_equal(lab, labs)


def test_full_sl_lifetime_value_errors():
    # all heights distinct: no lifetime close to zero -> numpy raises
    with pytest.raises(ValueError):
        O.full_sl_lifetime(np.array(GOLD["ytdist"]), 6)
    # all equal heights: one cluster, empty labels
    nc, lab, _ = O.full_sl_lifetime(np.ones(10), 5)
    assert nc == 1 and lab.size == 0
