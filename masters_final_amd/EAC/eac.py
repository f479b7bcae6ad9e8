"""Evidence-accumulation clustering driver pieces on the GPU, behind the
reference's API (MyML/MyML/EAC/eac.py).

sp_sl_lifetime replaces eac.py:199-267 (Kruskal minimum spanning tree,
lifetime rule of eac.py:366-391, cut, scipy connected components). The
forest is built by Boruvka on the GPU (eac_sl_lifetime, rule EAC_CUT_EAC);
the partition equals the reference's: with the lifetime rule the cut removes
exactly the forest edges heavier than the lower end of the largest gap,
which does not depend on which of several equal-weight spanning trees is
taken. Labels are numbered by first occurrence, as scipy's
connected_components numbers them.
"""
from __future__ import annotations

import numpy as np

from .. import _lib as L
from ..graph.singlelink import EAC_CUT_EAC, lifetime_cut


def _csr_parts(mat):
    """(indptr, indices, data) of a scipy CSR matrix or a tuple of arrays/tensors."""
    if isinstance(mat, tuple):
        return mat
    return mat.indptr, mat.indices, mat.data


def sp_sl_lifetime_device(mat, max_dist=None, n_clusters=0):
    """Device variant: (n_comps, labels tensor or None for one cluster, info)."""
    indptr, indices, data = _csr_parts(mat)
    labels, info = lifetime_cut(indices, data, indptr, None, rule=EAC_CUT_EAC,
                                max_dist=max_dist, n_clusters=n_clusters, symmetric=False)
    return info["n_components"], labels, info


def sp_sl_lifetime(mat, max_dist=None, n_clusters=0):
    """eac.py:199-267: (n_fclusts, labels) for a CSR co-association/dissimilarity
    graph (any orientation: every stored entry is an undirected edge). With one
    final cluster the labels are an empty int32 array, as in eac.py:263-265."""
    n_comps, labels, _ = sp_sl_lifetime_device(mat, max_dist, n_clusters)
    if labels is None:
        return 1, np.empty(0, dtype=np.int32)
    return n_comps, labels.cpu().numpy()
