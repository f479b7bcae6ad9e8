"""GPU K-Means labeling / centroid update / fit against the FP64 oracle."""
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


def _data(n, d, k, seed, spread=2.0):
    X, _ = synth.gmm_points(n, d, max(k // 2, 1), seed, mean_std=spread)
    rng = np.random.default_rng(seed + 100)
    Cc = X[rng.choice(n, k, replace=False)].copy()
    return X, Cc


@pytest.mark.parametrize("n,d,k", [(5000, 16, 37), (4000, 24, 100), (3000, 64, 256),
                                   (2000, 100, 64), (3000, 2, 75), (777, 33, 5),
                                   (3000, 2, 300), (2500, 1, 20), (2500, 3, 40),
                                   (2000, 5, 30)])
def test_label_bf16x3_equals_fp64_oracle(torch, n, d, k):
    from masters_final_amd.cluster.KMeans import label_device
    X, Cc = _data(n, d, k, seed=d * 7 + k)
    want, wd = O.kmeans_label_f64(X, Cc)
    lab, dist, nres = label_device(X, Cc, precision="bf16x3")
    assert np.array_equal(lab[0].cpu().numpy(), want)
    assert np.array_equal(dist[0].cpu().numpy(), wd.astype(np.float32))
    lab64, dist64, _ = label_device(X, Cc, precision="fp64")
    assert np.array_equal(lab64[0].cpu().numpy(), want)
    assert np.array_equal(dist64[0].cpu().numpy(), wd.astype(np.float32))


def test_label_ties_lowest_index(torch):
    from masters_final_amd.cluster.KMeans import label_device
    X, Cc = _data(3000, 32, 40, seed=5)
    Cd = np.concatenate([Cc[:10], Cc[:10], Cc[10:]])  # exact duplicates: ties
    want, _ = O.kmeans_label_f64(X, Cd)
    assert np.all(want < 10) or np.all((want < 10) | (want >= 20))
    for prec in ("bf16x3", "fp64"):
        lab, _, nres = label_device(X, Cd, precision=prec)
        assert np.array_equal(lab[0].cpu().numpy(), want)
    lab, _, nres = label_device(X, Cd, precision="bf16x3")
    assert nres > 0  # tied points cannot be certified: they are re-scored


# Point (X, X) against centroids (T0, T1) and (T1, T0): the differences are
# (a, b) and (b, a), so the oracle's separately rounded sums a^2 + b^2 are
# equal (a tie: the lower index wins), while a fused multiply-add over
# ascending dimensions rounds b^2 + fl(a^2) and a^2 + fl(b^2), and for these
# values makes the second centroid strictly closer. Found by an exact
# (fractions) search over X ~ 1e3, T ~ 1e-4.
FMA_SENSITIVE_TIES = [(680.3784790039062, -1.3656633200298529e-05, -3.7909856473561376e-05),
                      (463.11016845703125, 8.245135541073978e-05, -2.0252986359992065e-05),
                      (-1514.383544921875, 3.94981871068012e-05, -6.70565859763883e-05),
                      (-1193.2025146484375, -0.00014924639253877103, 3.663782763396739e-06)]


@pytest.mark.parametrize("d", [2, 16, 40])
@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_label_ties_need_separate_roundings(torch, d, prec):
    """Ties that exist only with separately rounded products and sums (the
    oracle's FP64 arithmetic) resolve to the lower index; a fused
    multiply-add would pick the other centroid. d = 16 and 40 send the
    split-BF16 path's tied points through the FP64 re-scoring (zero padding
    dimensions add exact zeros)."""
    from masters_final_amd.cluster.KMeans import label_device
    m = len(FMA_SENSITIVE_TIES)
    X = np.zeros((m, d), np.float32)
    C = np.zeros((2 * m, d), np.float32)
    for i, (x, t0, t1) in enumerate(FMA_SENSITIVE_TIES):
        X[i, :2] = x
        C[2 * i, :2] = (t0, t1)
        C[2 * i + 1, :2] = (t1, t0)
    lab, dist, _ = label_device(X, C, k_list=[2] * m, precision=prec)
    got = np.array([lab[i, i].item() for i in range(m)])
    want = np.array([O.kmeans_label_f64(X[i:i + 1], C[2 * i:2 * i + 2])[0][0] for i in range(m)])
    assert want.tolist() == [0] * m
    assert np.array_equal(got, want)


def test_rescore_many_flags_tiled(torch):
    """More than 65,536 flagged points (ties between exact duplicate
    centroids) take the partition-sorted, LDS-tiled FP64 re-scoring; labels
    must still be the oracle's (lowest index among equal distances)."""
    from masters_final_amd.cluster.KMeans import label_device
    rng = np.random.default_rng(31)
    X, _ = synth.gmm_points(70000, 16, 20, 31, mean_std=2.0)
    base0 = X[rng.choice(70000, 20, replace=False)]
    base1 = X[rng.choice(70000, 25, replace=False)]
    C0 = np.concatenate([base0, base0])                  # every point ties
    C1 = np.concatenate([base1[:5], base1, base1[5:9]])  # some points tie
    ks = [40, 34]
    lab, _, nres = label_device(X, np.concatenate([C0, C1]), k_list=ks, precision="bf16x3")
    assert nres >= 65536
    for p, Cp in enumerate((C0, C1)):
        want, _ = O.kmeans_label_f64(X, Cp)
        assert np.array_equal(lab[p].cpu().numpy(), want), p


@pytest.mark.parametrize("d", [16, 19, 40, 100])
def test_rescore_tiled_screen_near_ties(torch, d):
    """Far from the origin the split-BF16 bound flags nearly every point, so
    the tiled re-scoring decides (point coordinates in registers for d <= 64,
    centroid rows zero padded to 4 floats); centroids one ulp apart (and
    points on them) put FP32 screened values within the screen margin of
    each other. Labels must be the oracle's."""
    from masters_final_amd.cluster.KMeans import label_device
    rng = np.random.default_rng(37 + d)
    X, _ = synth.gmm_points(70000, d, 12, 37, mean_std=2.0)
    X = (X + 300.0).astype(np.float32)
    base = X[rng.choice(70000, 24, replace=False)]
    near = base.copy()
    j = rng.integers(0, d, 24)
    near[np.arange(24), j] = np.nextafter(near[np.arange(24), j], np.float32(np.inf))
    C0 = np.concatenate([base, near])                   # pairs one ulp apart
    C1 = np.concatenate([near[:9], base[3:20], near[9:]])
    X[:48] = np.concatenate([base, near])              # points exactly on centroids
    ks = [48, 41]
    lab, _, nres = label_device(X, np.concatenate([C0, C1]), k_list=ks, precision="bf16x3")
    assert nres >= 65536
    for p, Cp in enumerate((C0, C1)):
        want, _ = O.kmeans_label_f64(X, Cp)
        assert np.array_equal(lab[p].cpu().numpy(), want), p


def test_label_fp32_plain(torch):
    from masters_final_amd.cluster.KMeans import label_device
    X, Cc = _data(4000, 16, 50, seed=9)
    lab, dist, _ = label_device(X, Cc, precision="fp32")
    want, wd = O.kmeans_label_f64(X, Cc)
    agree = np.mean(lab[0].cpu().numpy() == want)
    assert agree > 0.999
    assert np.allclose(dist[0].cpu().numpy(), wd, rtol=1e-4, atol=1e-4)


def test_label_batched_partitions(torch):
    from masters_final_amd.cluster.KMeans import label_device
    rng = np.random.default_rng(3)
    X, _ = synth.gmm_points(6000, 16, 12, 3, mean_scale=20.0, anisotropic=(0.2, 3.0))
    ks = [32, 7, 300, 1, 64]
    C = np.concatenate([X[rng.choice(6000, k, replace=False)] for k in ks])
    lab, dist, _ = label_device(X, C, k_list=ks, precision="bf16x3")
    off = np.concatenate([[0], np.cumsum(ks)])
    for p, k in enumerate(ks):
        want, wd = O.kmeans_label_f64(X, C[off[p]:off[p + 1]])
        assert np.array_equal(lab[p].cpu().numpy(), want), p
        assert np.array_equal(dist[p].cpu().numpy(), wd.astype(np.float32)), p


def test_centroid_update_matches_oracle(torch):
    from masters_final_amd.cluster.KMeans import update_device
    rng = np.random.default_rng(4)
    X = rng.standard_normal((20000, 12)).astype(np.float32) * 3
    k = 40
    lab = rng.integers(0, k, 20000).astype(np.int32)
    lab[(lab == 7) | (lab == 21)] = 3  # two empty clusters
    dists = rng.random(20000).astype(np.float32)
    want = O.recompute_centroids(X, lab, dists, k, reseed_mode=1)
    got, ne = update_device(X, lab[None], dists[None], [k])
    assert ne == 2
    assert np.array_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize("d", [1, 3, 8, 16, 33, 64])
def test_centroid_update_narrow_rows_matches_oracle(torch, d):
    """d <= 64: several centroids per wavefront (4 to 64 lanes each); sizes
    from empty and single-member clusters to thousands of members (full
    batches of 16 rows plus a remainder), sums in point order."""
    from masters_final_amd.cluster.KMeans import update_device
    rng = np.random.default_rng(40 + d)
    n, k = 30000, 70
    X = (rng.standard_normal((n, d)) * 3 + 5).astype(np.float32)
    lab = np.minimum(rng.geometric(0.08, n) - 1, k - 1).astype(np.int32)  # skewed sizes
    lab[lab == 5] = 4                                                     # one empty cluster
    lab[:1] = 60                                                          # small clusters
    dists = rng.random(n).astype(np.float32)
    want = O.recompute_centroids(X, lab, dists, k, reseed_mode=1)
    got, ne = update_device(X, lab[None], dists[None], [k])
    assert ne == int(np.sum(np.bincount(lab, minlength=k) == 0))
    assert np.array_equal(got.cpu().numpy(), want)


def test_centroid_update_wide_rows_matches_oracle(torch):
    """More than 128 dimensions (two dimension groups) and clusters larger
    than one LDS stage of members."""
    from masters_final_amd.cluster.KMeans import update_device
    rng = np.random.default_rng(14)
    X = rng.standard_normal((3000, 130)).astype(np.float32)
    k = 6
    lab = rng.integers(0, k, 3000).astype(np.int32)
    dists = rng.random(3000).astype(np.float32)
    want = O.recompute_centroids(X, lab, dists, k, reseed_mode=1)
    got, ne = update_device(X, lab[None], dists[None], [k])
    assert ne == 0
    assert np.array_equal(got.cpu().numpy(), want)


def test_fit_batched_many_small_partitions_2d(torch):
    """The ensemble shape: 2-D points, many partitions of small k, partitions
    that stop at different iterations."""
    from masters_final_amd.cluster.KMeans import fit_batched
    rng = np.random.default_rng(15)
    X, _ = synth.gmm_points(3000, 2, 6, 15, mean_scale=10.0, std=1.0)
    ks = [int(k) for k in rng.integers(2, 60, 40)]
    init = np.concatenate([X[rng.choice(3000, k, replace=False)] for k in ks])
    r = fit_batched(X, init, ks, max_iter=6, tol=1e-2, precision="bf16x3")
    off = np.concatenate([[0], np.cumsum(ks)])
    for p, k in enumerate(ks):
        lab, dist, cen, it = O.kmeans_fit(X, init[off[p]:off[p + 1]], max_iter=6, tol=1e-2,
                                          reseed_mode=1)
        assert int(r["iters"][p]) == it, p
        assert np.array_equal(r["labels"][p].cpu().numpy(), lab), p
        assert np.array_equal(r["dists"][p].cpu().numpy(), dist), p
        assert np.array_equal(r["centroids"][off[p]:off[p + 1]].cpu().numpy(), cen), p


@pytest.mark.parametrize("prec", ["bf16x3", "fp64"])
def test_fit_matches_oracle(torch, prec):
    from masters_final_amd.cluster.KMeans import fit_batched
    X, Cc = _data(8000, 16, 60, seed=11)
    lab, dist, cen, it = O.kmeans_fit(X, Cc, max_iter=3, reseed_mode=1)
    r = fit_batched(X, Cc, [60], max_iter=3, precision=prec)
    assert int(r["iters"][0]) == it
    assert np.array_equal(r["labels"][0].cpu().numpy(), lab)
    assert np.array_equal(r["centroids"].cpu().numpy(), cen)
    assert np.array_equal(r["dists"][0].cpu().numpy(), dist)


def test_fit_batched_uneven_k(torch):
    from masters_final_amd.cluster.KMeans import fit_batched
    rng = np.random.default_rng(12)
    X, _ = synth.gmm_points(10000, 16, 32, 12, mean_scale=20.0, anisotropic=(0.2, 3.0))
    ks = [32, 150, 45, 512, 33]
    init = np.concatenate([X[rng.choice(10000, k, replace=False)] for k in ks])
    r = fit_batched(X, init, ks, max_iter=3, precision="bf16x3")
    off = np.concatenate([[0], np.cumsum(ks)])
    for p, k in enumerate(ks):
        lab, dist, cen, it = O.kmeans_fit(X, init[off[p]:off[p + 1]], max_iter=3, reseed_mode=1)
        assert np.array_equal(r["labels"][p].cpu().numpy(), lab), p
        assert np.array_equal(r["centroids"][off[p]:off[p + 1]].cpu().numpy(), cen), p


def test_k_means_class(torch):
    from masters_final_amd.cluster.KMeans import K_Means
    X, y = synth.eac_small(seed=0)
    init = X[:70].copy()
    km = K_Means(n_clusters=70, max_iter=3, init=init)
    km.fit(X)
    lab, dist, cen, it = O.kmeans_fit(X, init, max_iter=3, reseed_mode=1)
    assert km.iters_ == it and np.array_equal(km.labels_, lab)
    with pytest.raises(Warning):
        K_Means(n_clusters=3).fit(X.astype(np.float64))


def test_sum_f32_rows_equal_numpy(torch):
    """eac_sum_f32_rows gives numpy's float32 sum to the bit (the inertia of
    KMeans.py:164): row lengths around the pairwise block and the 8192-element
    reduction buffer, magnitudes over 40 binades, many rows at once."""
    from masters_final_amd.cluster.KMeans import sum_f32_device
    rng = np.random.default_rng(21)
    for n in [1, 7, 8, 9, 127, 128, 129, 1000, 8191, 8192, 8193, 10000, 65537, 500_000]:
        rows = 3 if n > 100_000 else 17
        x = (rng.standard_normal((rows, n)) * np.exp(rng.uniform(-20, 20, (rows, n)))).astype(
            np.float32)
        got = sum_f32_device(x).cpu().numpy()
        want = np.array([x[r].sum() for r in range(rows)], np.float32)
        assert got.tobytes() == want.tobytes(), n
    big = rng.random(4_194_304).astype(np.float32)
    assert sum_f32_device(big).cpu().numpy()[0] == big.sum()


def test_fit_float32_stop_rule(torch):
    """Two labelings: the second changes the dists by 0.013 but not their
    float32 sum (~2e8), so the reference stops where float64 sums would not."""
    from masters_final_amd.cluster.KMeans import K_Means, fit_batched
    X = np.array([[-1.0], [0.0], [1.2], [20000.0], [40000.0]], np.float32)
    init = np.array([[0.0], [30000.0]], np.float32)
    lab, dist, cen, it, inert = O.kmeans_fit(X, init, max_iter=3, tol=1e-4, reseed_mode=1,
                                             return_inertia=True)
    assert it == 2
    r = fit_batched(X, init, [2], max_iter=3, precision="fp64")
    assert int(r["iters"][0]) == 2
    assert np.array_equal(r["labels"][0].cpu().numpy(), lab)
    assert np.array_equal(r["centroids"].cpu().numpy(), cen)
    km = K_Means(n_clusters=2, max_iter=3, init=init)
    km.fit(X)
    assert km.iters_ == 2 and km.inertia_.tobytes() == inert.tobytes()


@pytest.mark.parametrize("d,k_list", [(20, None), (64, None), (100, None),
                                      (64, [1, 63, 64, 65, 150, 7, 200]), (16, [33, 2, 129])])
def test_label_waves_knob_keeps_labels(torch, d, k_list):
    """The waves-per-SIMD variants of the split-BF16 labeling kernel
    (km_waves 2, 3, 4, built-in) give the same labels and dists, for one
    centroid set and for several (partition sizes below, at and past the
    64-centroid tile)."""
    from masters_final_amd.cluster.KMeans import label_device
    from masters_final_amd.tune import get_knobs, set_knobs
    rng = np.random.default_rng(d)
    X = rng.normal(size=(3000, d)).astype(np.float32)
    k = 150 if k_list is None else sum(k_list)
    Cc = X[rng.choice(3000, k, replace=False)].copy()
    dX, dC = torch.from_numpy(X).cuda(), torch.from_numpy(Cc).cuda()
    saved = get_knobs()
    outs = []
    try:
        for w in (0, 2, 3, 4):
            set_knobs(km_waves=w)
            lab, dist, _ = label_device(dX, dC, k_list=k_list, precision="bf16x3")
            outs.append((lab.clone(), dist.clone()))
    finally:
        set_knobs(**saved)
    for lab, dist in outs[1:]:
        assert torch.equal(lab, outs[0][0]) and torch.equal(dist, outs[0][1])
