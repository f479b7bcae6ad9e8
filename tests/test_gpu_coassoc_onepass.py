"""The one-pass co-association build (sampled nnz estimate, staged pass into
pool chunks, compaction) gives the two-pass build's matrix and the oracle's
(EAC_CSR exact condensed or full rows, MyML/MyML/EAC/sparse.py:154-225,
make_diassoc :295-311); a pool too small for the staged pass is detected and
the build falls back to two passes with the same result."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def _labels(n, P, kmin, kmax, seed, unassigned=0):
    rng = np.random.default_rng(seed)
    ks = rng.integers(kmin, kmax, P)
    lab = np.stack([rng.integers(0, k, n) for k in ks]).astype(np.int32)
    if unassigned:
        lab[rng.integers(0, P, unassigned), rng.integers(0, n, unassigned)] = -1
    return lab, ks


def _build(labels, ks, **kw):
    from masters_final_amd.EAC.sparse import coassoc_device
    info = {}
    ip, ind, dat, mx = coassoc_device(labels, ks, info=info, **kw)
    return ip.cpu().numpy(), ind.cpu().numpy(), dat.cpu().numpy(), mx, info


@pytest.mark.parametrize("n,P,kmin,kmax,window,stride,seed", [
    (3000, 24, 5, 40, 64, 7, 1),       # 47 windows per row, many pool chunks
    (9000, 16, 20, 90, 1000, 32, 2),   # uneven last window
    (5000, 40, 2, 12, 0, 1, 3),        # default window (the whole row range), exact estimate
    (1, 3, 1, 3, 0, 1, 4),             # one point: empty matrix
])
def test_one_pass_condensed_equals_oracle(torch, n, P, kmin, kmax, window, stride, seed):
    from masters_final_amd.tune import get_knobs, set_knobs
    labels, ks = _labels(n, P, kmin, kmax, seed, unassigned=n // 10)
    saved = get_knobs()
    try:
        set_knobs(co_window=window)
        gi, gc, gd, mx, info = _build(labels, ks, one_pass=True, sample_stride=stride)
    finally:
        set_knobs(**saved)
    assert info["path"] == "one-pass"
    ip, ind, cnt = O.coassoc_condensed(labels, ks)
    assert np.array_equal(gi, ip) and np.array_equal(gc, ind) and np.array_equal(gd, cnt)
    assert mx == (int(cnt.max()) if cnt.size else 0)


@pytest.mark.parametrize("ranges", [1, 3, 8])
def test_one_pass_row_ranges_keep_the_matrix(torch, ranges):
    """Any number of row ranges (the default is one below 65536 rows, one
    per XCD above) gives the oracle's matrix, one-pass and two-pass."""
    from masters_final_amd.tune import get_knobs, set_knobs
    labels, ks = _labels(6000, 30, 4, 60, 11, unassigned=500)
    saved = get_knobs()
    try:
        set_knobs(co_ranges=ranges)
        a = _build(labels, ks, one_pass=True)
        b = _build(labels, ks, one_pass=False)
    finally:
        set_knobs(**saved)
    ip, ind, cnt = O.coassoc_condensed(labels, ks)
    for gi, gc, gd in (a[:3], b[:3]):
        assert np.array_equal(gi, ip) and np.array_equal(gc, ind) and np.array_equal(gd, cnt)


def test_one_pass_full_rows_dissim_equals_two_pass(torch):
    from masters_final_amd.tune import get_knobs, set_knobs
    labels, ks = _labels(2500, 20, 3, 30, 5, unassigned=300)
    saved = get_knobs()
    try:
        set_knobs(co_window=128)
        a = _build(labels, ks, full_rows=True, w_offset="max+1", one_pass=True, sample_stride=5)
        b = _build(labels, ks, full_rows=True, w_offset="max+1", one_pass=False)
    finally:
        set_knobs(**saved)
    assert a[4]["path"] == "one-pass" and b[4]["path"] == "two-pass"
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    # full rows are symmetric and hold (max + 1) - count
    gi, gc, gd, mx = a[:4]
    rows = np.repeat(np.arange(gi.size - 1), np.diff(gi))
    D = {(int(r), int(c)): int(v) for r, c, v in zip(rows, gc, gd)}
    assert all(D[(c, r)] == v for (r, c), v in D.items())
    assert gd.min() >= 1 and gd.max() <= mx


def test_pool_overflow_falls_back_to_two_passes(torch):
    """A pool far below the estimate: the staged pass reports the overflow
    and the build runs the two passes; the same matrix either way."""
    from masters_final_amd.tune import get_knobs, set_knobs
    labels, ks = _labels(4000, 30, 4, 25, 6)
    saved = get_knobs()
    try:
        set_knobs(co_window=256)
        a = _build(labels, ks, one_pass=True, sample_stride=3, margin=-0.95)
        b = _build(labels, ks, one_pass=False)
        # and a second one-pass build after the overflow: clean state
        c = _build(labels, ks, one_pass=True, sample_stride=3, margin=0.5)
    finally:
        set_knobs(**saved)
    assert a[4]["path"] == "one-pass overflow, two-pass"
    assert c[4]["path"] == "one-pass"
    for x, y, z in zip(a[:4], b[:4], c[:4]):
        assert np.array_equal(x, y) and np.array_equal(x, z)


def test_estimate_is_exact_with_stride_one(torch):
    from masters_final_amd import _lib as L
    import ctypes as C
    labels, ks = _labels(1500, 12, 3, 20, 7)
    gi = _build(labels, ks, one_pass=False)[0]
    lib = L.load_library()
    dl = torch.from_numpy(labels).cuda()
    koff = np.zeros(len(ks) + 1, np.int64)
    np.cumsum(ks, out=koff[1:])
    ws = L.workspace(lib.eac_coassoc_workspace_bytes(1500, 12, int(koff[-1])))
    for stride, exact in ((1, True), (4, False)):
        est = C.c_int64()
        L.check(lib.eac_coassoc_estimate(L.ptr(dl), 1500, koff.ctypes.data_as(C.c_void_p), 12, 0,
                                         stride, C.byref(est), L.ptr(ws), ws.numel(),
                                         L.stream_ptr()))
        if exact:
            assert est.value == gi[-1]
        else:
            assert abs(est.value - gi[-1]) <= 0.2 * gi[-1]
    assert lib.eac_coassoc_estimate(L.ptr(dl), 1500, koff.ctypes.data_as(C.c_void_p), 12, 0, 0,
                                    C.byref(est), L.ptr(ws), ws.numel(),
                                    L.stream_ptr()) == L.EAC_ERR_ARG
