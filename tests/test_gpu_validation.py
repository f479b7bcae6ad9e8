"""GPU contingency table and the accuracy indices built on it, against numpy
and the literal restatement of the reference."""
import numpy as np
import pytest

from oracle.validation_oracle import consistency_literal, hungarian_bruteforce

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def _np_table(a, b, ka, kb, a0=0, b0=0):
    T = np.zeros((ka, kb), np.int64)
    ok = (a - a0 >= 0) & (a - a0 < ka) & (b - b0 >= 0) & (b - b0 < kb)
    np.add.at(T, (a[ok] - a0, b[ok] - b0), 1)
    return T, int((~ok).sum())


@pytest.mark.parametrize("n,ka,kb", [(0, 3, 4), (1, 1, 1), (10000, 7, 13), (300000, 64, 128),
                                     (200000, 100, 100), (50000, 300, 400)])
def test_contingency_kernel(torch, n, ka, kb):
    """LDS tables (ka*kb <= 8192) and global-atomic tables; labels outside
    the range are counted in the extra slot."""
    from masters_final_amd import _lib as L
    rng = np.random.default_rng(n + ka)
    a = rng.integers(-1, ka + 1, n).astype(np.int32)
    b = rng.integers(0, kb, n).astype(np.int32)
    lib = L.load_library()
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = torch.full((ka * kb + 1,), -5, dtype=torch.int64, device="cuda")
    L.check(lib.eac_contingency(L.ptr(da), L.ptr(db), n, 0, ka, 0, kb, L.ptr(out),
                                L.stream_ptr()))
    h = out.cpu().numpy()
    T, bad = _np_table(a, b, ka, kb)
    assert np.array_equal(h[:-1].reshape(ka, kb), T)
    assert h[-1] == bad


def test_indices_match_reference(torch):
    from masters_final_amd.validation.consistency import ConsistencyIndex
    from masters_final_amd.validation.hindex import HungarianIndex
    rng = np.random.default_rng(5)
    for trial in range(12):
        n = int(rng.integers(20, 200))
        a = np.unique(rng.integers(0, int(rng.integers(2, 7)), n), return_inverse=True)[1]
        b = np.unique(rng.integers(0, int(rng.integers(2, 7)), n), return_inverse=True)[1]
        if trial % 3 == 0:
            b = b + 1
        want, acc = consistency_literal(a, b, n)
        ci = ConsistencyIndex(N=n)
        assert ci.score(a, b) == acc and ci.match_count == want
        T, _ = _np_table(a, b - b.min(), a.max() + 1, b.max() - b.min() + 1)
        hi = HungarianIndex(n)
        assert hi.score(a, b) == hungarian_bruteforce(T) / n
        # list and bin formats describe the same partitions
        la = [np.nonzero(a == c)[0] for c in range(a.max() + 1)]
        lb = [np.nonzero(b == c)[0] for c in range(b.min(), b.max() + 1)]
        assert ConsistencyIndex(N=n).score(la, lb, format='list') == acc
        assert ConsistencyIndex(N=n).score(a, lb, format1='array', format2='list') == acc
        binb = np.zeros((len(lb), n), np.uint8)
        for c, idx in enumerate(lb):
            binb[c, idx] = 1
        assert HungarianIndex(n).score(la, binb, format1='list', format2='bin') == hi.accuracy


def test_indices_large_partitions(torch):
    """Overlaps far above 255 (the reference's uint8 matrices would wrap):
    a relabelled copy scores 1, an ensemble partition against ground truth is
    bounded by the optimum."""
    from masters_final_amd.validation.consistency import ConsistencyIndex
    from masters_final_amd.validation.hindex import HungarianIndex
    rng = np.random.default_rng(9)
    n = 200000
    a = rng.integers(0, 40, n)
    perm = rng.permutation(40)
    assert ConsistencyIndex(N=n).score(a, perm[a]) == 1.0
    assert HungarianIndex(n).score(a, perm[a]) == 1.0
    b = np.where(rng.random(n) < 0.9, a, rng.integers(0, 40, n))
    c = ConsistencyIndex(N=n).score(a, b)
    h = HungarianIndex(n).score(a, b)
    assert 0.9 <= c <= h <= 1.0


def test_indices_errors(torch):
    from masters_final_amd.validation.consistency import ConsistencyIndex
    from masters_final_amd.validation.hindex import HungarianIndex
    a = np.array([0, 0, 1, 1])
    with pytest.raises(ValueError):
        ConsistencyIndex(N=4).score(a, np.array([0, 3, 3, 5]))  # gaps in the numbering
    with pytest.raises(Exception):
        ConsistencyIndex(N=4).score(a, a, format='bin')
    with pytest.raises(ValueError):
        HungarianIndex(4).score(a, a, format='dict')
    # {0, 2}: passes the reference's count check, but a row has no sample
    with pytest.raises(ValueError):
        ConsistencyIndex(N=4).score(np.array([0, 0, 2, 2]), np.array([0, 1, 2, 2]))
    # a huge label is rejected before any table is allocated
    with pytest.raises(ValueError):
        ConsistencyIndex(N=2).score(np.array([0, 10 ** 9]), np.array([0, 1]))
    # 'list' format: one sample listed in two clusters
    with pytest.raises(ValueError):
        HungarianIndex(4).score([np.array([0, 1]), np.array([1, 2, 3])],
                                [np.array([0, 1]), np.array([2, 3])], format='list')


def test_bin_format_with_an_empty_column(torch):
    """A binary partition whose last sample is in no cluster: N is the matrix
    width, the sample counts in no cluster (the 'bin' path used to size the
    labels from the number of ones and fail)."""
    from masters_final_amd.validation.contingency import contingency
    b1 = np.array([[1, 1, 0, 0], [0, 0, 1, 0]], np.uint8)
    b2 = np.array([[1, 0, 1, 0], [0, 1, 0, 0]], np.uint8)
    T, s1, s2, n = contingency(b1, b2, format1='bin', format2='bin')
    assert n == 4 and T.tolist() == [[1, 1], [1, 0]]
    assert s1.tolist() == [2, 1] and s2.tolist() == [2, 1]
