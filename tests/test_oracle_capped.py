"""Pins the oracle's capacity-capped co-association (orc_coassoc_capped, the
restatement of the reference's surgical EAC_CSR build with max_assocs,
MyML/MyML/EAC/sparse.py:154-225, :368-465, :866-1099).

Checks, each independent of the restated insertion loop:
  * with capacities no row can reach, the result is the exact co-association
    (the uncapped oracle, itself pinned by a dense brute force);
  * a set formulation of the cap: a row keeps the first `cap` of its
    co-occurring columns in the order (first partition, in processing order,
    in which the pair co-occurs; column), every kept column carries its full
    count, and the discard count is the sum of the dropped columns' counts;
  * hand cases: capacity 1, the first-partition overflow, the full-row
    debug return above id 405109, empty labels.
"""
import numpy as np
import pytest

from masters_final_amd.EAC.sparse import indptr_constant, indptr_linear
from oracle import oracle as O


def _processing_order(L):
    """sparse.py:154-166: fewest distinct labels first (first minimum), then
    the others in order."""
    P = L.shape[0]
    nd = [np.unique(L[p][L[p] >= 0]).size for p in range(P)]
    f = int(np.argmin(nd))
    return [f] + [p for p in range(P) if p != f]


def _set_model(L, cap_off, full_rows=False, max_assocs=0):
    P, n = L.shape
    order = _processing_order(L)
    rows_c, rows_v, disc = [], [], 0
    for i in range(n):
        first, cnt = {}, {}
        for rank, p in enumerate(order):
            li = L[p, i]
            if li < 0:
                continue
            if full_rows and rank > 0 and i > 405109:
                continue
            js = np.nonzero(L[p] == li)[0]
            js = js[js != i] if full_rows else js[js > i]
            for j in js:
                cnt[j] = cnt.get(j, 0) + 1
                first.setdefault(j, rank)
        cap = max_assocs if full_rows else cap_off[i + 1] - cap_off[i]
        n_first = sum(1 for r in first.values() if r == 0)
        if n_first > cap:
            raise ValueError(i)
        ranked = sorted(first, key=lambda j: (first[j], j))
        keep = sorted(ranked[:cap])
        disc += sum(cnt[j] for j in ranked[cap:])
        rows_c.append(keep)
        rows_v.append([cnt[j] for j in keep])
    indptr = np.concatenate([[0], np.cumsum([len(r) for r in rows_c])])
    return (indptr, np.array(sum(rows_c, []), np.int64), np.array(sum(rows_v, []), np.int64),
            disc)


def _labels(rng, n, P, kmin, kmax):
    return np.stack([rng.integers(0, rng.integers(kmin, kmax + 1), n) for _ in range(P)])


def test_capped_unreachable_caps_equal_exact():
    rng = np.random.default_rng(0)
    n, P = 300, 12
    L = _labels(rng, n, P, 3, 9)
    ks = L.max(1) + 1
    ip, ind, cnt = O.coassoc_condensed(L, ks)
    off = indptr_constant(n, n)
    cip, cind, ccnt, disc = O.coassoc_capped(L, off)
    assert disc == 0
    assert np.array_equal(cip, ip) and np.array_equal(cind, ind) and np.array_equal(ccnt, cnt)


@pytest.mark.parametrize("seed,n,P,kmin,kmax,fac", [(1, 60, 6, 2, 5, 3), (2, 80, 8, 3, 8, 3),
                                                    (3, 50, 10, 2, 4, 1.5), (4, 70, 5, 4, 9, 2),
                                                    (7, 120, 12, 5, 12, 1)])
def test_capped_linear_equals_set_model(seed, n, P, kmin, kmax, fac):
    """max_assocs = fac x the biggest cluster (eac.py:118-120 uses 3): no
    discards at 3, discards at 1.5 and 2, a first-partition overflow at 1."""
    rng = np.random.default_rng(seed)
    L = _labels(rng, n, P, kmin, kmax)
    ma = int(fac * max(np.bincount(lab).max() for lab in L))
    off = indptr_linear(n, ma, 0.05, 1.0, 1.0, 0.05)
    try:
        want = _set_model(L, off)
    except ValueError:
        with pytest.raises(ValueError):
            O.coassoc_capped(L, off)
        return
    got = O.coassoc_capped(L, off)
    assert got[3] == want[3]
    for a, b in zip(got[:3], want[:3]):
        assert np.array_equal(a, b)


def test_capped_discards_happen():
    """Small constant caps after a first partition that fits: discards are
    counted and the set model agrees."""
    rng = np.random.default_rng(5)
    n, P = 64, 9
    L = _labels(rng, n, P, 6, 10)
    L[0] = np.arange(n) % 16  # 16 clusters of 4
    L[1] = np.arange(n) % 4   # 4 clusters of 16: the fewest, so processed first
    off = indptr_constant(n, 18)
    want = _set_model(L, off)
    got = O.coassoc_capped(L, off)
    assert want[3] > 0 and got[3] == want[3]
    for a, b in zip(got[:3], want[:3]):
        assert np.array_equal(a, b)


def test_capped_full_rows_constant():
    rng = np.random.default_rng(6)
    n, P = 50, 7
    L = _labels(rng, n, P, 3, 6)
    L[2] = np.arange(n) % 3  # first partition: 3 clusters of <= 17 members
    ma = 20
    off = indptr_constant(n, ma)
    want = _set_model(L, off, full_rows=True, max_assocs=ma)
    got = O.coassoc_capped(L, off, full_rows=True, max_assocs=ma)
    assert got[3] == want[3]
    for a, b in zip(got[:3], want[:3]):
        assert np.array_equal(a, b)


def test_capped_hand_case_capacity_one():
    # two partitions over 4 points; first processed: p1 (1 cluster)
    L = np.array([[0, 0, 1, 1],
                  [0, 0, 0, 0]])
    off = np.array([0, 3, 5, 6, 6])  # capacities 3, 2, 1, 0
    ip, ind, cnt, disc = O.coassoc_capped(L, off)
    # p1 first: row 0 -> {1,2,3}, row 1 -> {2,3}, row 2 -> {3}; p0 adds 1 to
    # (0,1) and (2,3): all found, nothing new
    assert ip.tolist() == [0, 3, 5, 6, 6]
    assert ind.tolist() == [1, 2, 3, 2, 3, 3] and cnt.tolist() == [2, 1, 1, 1, 1, 2]
    assert disc == 0
    # p0 (2 clusters) goes before p1 (3 clusters); capacity 1 per row
    L2 = np.array([[0, 0, 1, 1],
                   [0, 1, 1, 2]])
    off2 = np.array([0, 1, 2, 3, 3])
    ip, ind, cnt, disc = O.coassoc_capped(L2, off2)
    # p0 fills row 0 with (0,1) and row 2 with (2,3); row 1 has no later
    # co-member in p0, so p1's pair (1,2) is its first entry
    assert ip.tolist() == [0, 1, 2, 3, 3]
    assert ind.tolist() == [1, 2, 3] and cnt.tolist() == [1, 1, 1]
    assert disc == 0
    # capacity 1 and a second new column: discarded, and counted each time
    L3 = np.array([[0, 0, 1, 1],
                   [0, 1, 0, 1],
                   [0, 1, 0, 1]])
    ip, ind, cnt, disc = O.coassoc_capped(L3, np.array([0, 1, 2, 3, 3]))
    # p0 first (ties on 2 clusters: the first minimum): row 0 keeps (0,1);
    # p1 and p2 both pair 0 with 2 (discarded twice) and 1 with 3 (row 1's
    # first entry, then +1)
    assert ind.tolist() == [1, 3, 3] and cnt.tolist() == [1, 2, 1]
    assert disc == 2


def test_capped_first_partition_overflow_raises():
    L = np.array([[0, 0, 0, 0]])
    with pytest.raises(ValueError, match="row 0"):
        O.coassoc_capped(L, indptr_constant(4, 2))


def test_capped_full_rows_debug_return():
    """Full rows: ids above 405109 take only the first partition's entries
    (sparse.py:877-878); the set model says so too."""
    n = 405113
    L = np.full((2, n), -1, np.int64)
    ids = np.array([405100, 405108, 405110, 405112])
    L[0, ids] = [0, 0, 0, 1]   # 2 clusters: first processed
    L[1, ids] = [0, 0, 0, 0]   # 1 cluster... fewer: this one goes first
    ma = 3
    off = indptr_constant(n, ma)
    ip, ind, cnt, disc = O.coassoc_capped(L, off, full_rows=True, max_assocs=ma)
    rows = np.repeat(np.arange(n), np.diff(ip))
    got = {(int(r), int(c)): int(v) for r, c, v in zip(rows, ind, cnt)}
    # first partition (L[1]): all four together; then L[0]'s cluster
    # {405100, 405108, 405110}: rows 405100 and 405108 add +1 to their two
    # co-members, row 405110 returns
    assert got[(405100, 405108)] == 2 and got[(405108, 405110)] == 2
    assert got[(405110, 405100)] == 1 and got[(405110, 405108)] == 1
    assert got[(405112, 405100)] == 1
    assert disc == 0
