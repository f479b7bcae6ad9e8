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
