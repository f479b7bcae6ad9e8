"""The all-core CPU baseline (orc_boruvka_omp, OpenMP Borůvka) computes the
canonical forest: the same edge set as the canonical Kruskal restatement
and the same weight as scipy's minimum spanning tree."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import minimum_spanning_tree

from masters_final_amd import synth
from oracle import oracle as O


@pytest.mark.parametrize("n,deg,seed,wmax", [(3000, 4, 1, 9), (20000, 6, 2, 200), (500, 30, 3, 3)])
def test_boruvka_omp_equals_canonical_kruskal(n, deg, seed, wmax):
    g = synth.random_graph(n, deg, seed, wmax=wmax, n_isolated=7, n_parts=3)
    fe, od = g["indptr"][:-1], np.diff(g["indptr"])
    u, v, w, _ = O.csr_edges(g["dest"], g["weight"], fe, od)
    inf, _ = O.kruskal_canonical(u, v, w, g["n"])
    lo, hi, tw, threads = O.boruvka_omp(g["indptr"], g["dest"], g["weight"].astype(np.int32))
    assert threads >= 1
    assert set(zip(lo.tolist(), hi.tolist())) == {(a, b) for a, b, f in zip(u, v, inf) if f}
    m = sp.csr_matrix((w.astype(np.float64), (u, v)), shape=(g["n"], g["n"]))
    assert tw == int(w[inf].sum()) == int(minimum_spanning_tree(m).sum())


def test_sl_batched_omp_equals_per_graph_oracle():
    """The all-core batched cut baseline gives every graph the per-graph
    oracle's components and labels."""
    gs = [synth.random_graph(n, deg, s, wmax=w, n_isolated=2, n_parts=p)
          for n, deg, s, w, p in [(400, 4, 5, 6, 2), (50, 3, 6, 3, 1), (1200, 6, 7, 30, 3),
                                  (2, 1, 8, 1, 1), (700, 5, 9, 1, 1)]]
    goff = np.concatenate([[0], np.cumsum([g["n"] for g in gs])]).astype(np.int64)
    eoff = np.concatenate([[0], np.cumsum([g["dest"].size for g in gs])]).astype(np.int64)
    indptr = np.concatenate([g["indptr"][:-1] + eoff[i] for i, g in enumerate(gs)] + [[eoff[-1]]])
    dest = np.concatenate([g["dest"] + goff[i] for i, g in enumerate(gs)]).astype(np.int32)
    w = np.concatenate([g["weight"] for g in gs]).astype(np.int32)
    ncs, labs, _ = O.sl_batched_omp(indptr, dest, w, goff)
    for i, g in enumerate(gs):
        fe, od = g["indptr"][:-1], np.diff(g["indptr"])
        u, v, ww, _ = O.csr_edges(g["dest"], g["weight"], fe, od)
        try:
            nc, want = O.sp_sl_lifetime(u, v, ww, g["n"])
        except ValueError:
            assert ncs[i] == -1
            continue
        if want.size == 0:
            assert ncs[i] == 0
            continue
        assert ncs[i] == nc and np.array_equal(labs[goff[i]:goff[i + 1]], want)
