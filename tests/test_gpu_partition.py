"""Partition formats, the ensemble generator, outdegree_from_firstedge and
np_label on the GPU.

Golden vectors: tests/golden/partition_formats.json holds the known answers
of the reference's own tests (MyML/tests/MyML/utils/test_partition.py:5-97)
for convertIndexToBin, convertClusterStringToBin and
convertClusterStringToIndex.
"""
import json
import random
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "partition_formats.json").read_text())


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def test_convert_index_to_bin_golden(torch):
    from masters_final_amd.utils.partition import convertIndexToBin
    g = GOLD["index_to_bin_small"]
    out = convertIndexToBin([np.array(c, np.uint64) for c in g["clusters"]])
    assert out.dtype == np.uint8
    assert np.array_equal(out, np.array(g["bin"], np.uint8))


def test_convert_cluster_string_to_index_golden(torch):
    from masters_final_amd.utils.partition import convertClusterStringToIndex
    g = GOLD["string_to_index_small"]
    out = convertClusterStringToIndex(np.array(g["labels"], np.uint32))
    assert [a.tolist() for a in out] == g["index"]
    assert all(a.dtype == np.uint64 for a in out)


def test_convert_index_to_bin_large(torch):
    """test_partition.py:20-43 pattern: random clusters of 1-199 samples
    covering 100,000 samples; every column holds exactly one 1, in the row of
    the cluster that lists it."""
    from masters_final_amd.utils.partition import convertIndexToBin
    rng = np.random.default_rng(7)
    n = 100_000
    perm = rng.permutation(n)
    cuts = np.cumsum(rng.integers(1, 200, n))
    cuts = cuts[cuts < n]
    clusts = [c.astype(np.uint64) for c in np.split(perm, cuts)]
    out = convertIndexToBin(clusts)
    assert out.shape == (len(clusts), n)
    assert np.array_equal(out.sum(0), np.ones(n))
    owner = np.repeat(np.arange(len(clusts)), [c.size for c in clusts])
    assert np.all(out[owner, perm] == 1)
    assert int(out.sum()) == n


@pytest.mark.parametrize("case", ["string_to_bin_min0", "string_to_bin_min1"])
def test_convert_cluster_string_to_bin_golden(torch, case):
    from masters_final_amd.utils.partition import convertClusterStringToBin
    g = GOLD[case]
    out = convertClusterStringToBin(np.array(g["labels"], np.uint64))
    assert np.array_equal(out, np.array(g["bin"], np.uint8))


def test_convert_cluster_string_to_bin_errors(torch):
    from masters_final_amd.utils.partition import convertClusterStringToBin
    with pytest.raises(ValueError):
        convertClusterStringToBin(np.array([0, 3, 3, 5]))  # 3 values, max 5
    # {0, 2}: the count check passes (2 values, max 2), the assignment is out
    # of range in the reference (IndexError)
    with pytest.raises(IndexError):
        convertClusterStringToBin(np.array([0, 0, 2, 2]))


def test_outdegree_from_firstedge(torch, elt4):
    """build.py:275-280 (test_boruvka.py:37-42): out-degrees from the 4elt
    offsets equal the fixture's own row lengths, on host and device arrays."""
    from masters_final_amd.graph.build import outdegree_from_firstedge
    fe = elt4["fe"].astype(np.int32)
    m = elt4["dest"].size
    od = np.zeros_like(fe)
    outdegree_from_firstedge(fe, od, m)
    assert np.array_equal(od, elt4["od"])
    assert np.array_equal(od, O.outdegree_from_firstedge(fe, m))
    dod = torch.zeros(fe.size, dtype=torch.int32, device="cuda")
    outdegree_from_firstedge(torch.from_numpy(fe).cuda(), dod, m)
    assert np.array_equal(dod.cpu().numpy(), elt4["od"])
    # a trailing vertex with no edges
    fe2 = np.array([0, 2, 2, 5], np.int64)
    od2 = np.zeros(4, np.int64)
    outdegree_from_firstedge(fe2, od2, 5)
    assert od2.tolist() == [2, 0, 3, 0]


def _np_label(X, Cc):
    """KMeans.py:493-522 as written (numpy float32 arithmetic)."""
    labels = np.zeros(X.shape[0], np.int32)
    best = ((X - Cc[0]) ** 2).sum(axis=1)
    for c in range(1, Cc.shape[0]):
        dist = ((X - Cc[c]) ** 2).sum(axis=1)
        labels[dist < best] = c
        best = np.minimum(dist, best)
    return labels, best


@pytest.mark.parametrize("d", [1, 2, 3, 7, 8, 9, 16, 31, 64, 100, 128])
def test_np_label_equals_numpy(torch, d):
    from masters_final_amd.cluster.KMeans import label_device, np_label
    rng = np.random.default_rng(100 + d)
    n, k = 3000, 37
    X = (rng.standard_normal((n, d)) * rng.uniform(0.5, 50.0, d)).astype(np.float32)
    Cc = X[rng.choice(n, k, replace=False)].copy()
    Cc[5] = Cc[3]  # exact duplicate centroid: the lower index must win
    want, wd = _np_label(X, Cc)
    dists = np.full(n, -1.0, np.float32)
    got = np_label(X, Cc, dists, True)
    assert np.array_equal(got, want)
    assert np.all(dists == -1.0)  # the reference never writes them back
    lab, dd, _ = label_device(X, Cc, precision="numpy32")
    assert np.array_equal(lab[0].cpu().numpy(), want)
    assert np.array_equal(dd[0].cpu().numpy(), wd)


def test_generate_ensemble_random_init_equals_oracle(torch):
    """partition.py:88-124 with K_Means 'random' init (KMeans.py:191-201):
    the same np.random / random streams drive k and the initial samples, the
    oracle fit gives the expected labels, converted to index lists as
    partition.py:73-86 defines them (ascending labels, members ascending)."""
    from masters_final_amd import synth
    from masters_final_amd.cluster.KMeans import K_Means
    from masters_final_amd.utils.partition import generateEnsemble
    X, _ = synth.gmm_points(1500, 2, 6, 9, mean_scale=10.0)
    np.random.seed(11)
    random.seed(12)
    gen = K_Means(n_clusters=5, precision="fp64")
    ens = generateEnsemble(X, gen, n_clusters=[5, 12], npartitions=6, iters=3)
    np.random.seed(11)
    random.seed(12)
    for part in ens:
        k = np.random.randint(5, 12)
        idx = random.sample(range(X.shape[0]), k)
        lab = O.kmeans_fit(X, X[idx], max_iter=3, reseed_mode=1)[0]
        want = [np.nonzero(lab == c)[0] for c in np.unique(lab)]
        assert len(part) == len(want)
        for a, b in zip(part, want):
            assert a.dtype == np.uint64 and np.array_equal(a, b)
    # fixed k (n_clusters an int)
    gen2 = K_Means(n_clusters=4, precision="bf16x3")
    ens2 = generateEnsemble(X, gen2, n_clusters=4, npartitions=2, iters=2)
    assert all(len(p) <= 4 and sum(a.size for a in p) == X.shape[0] for p in ens2)
