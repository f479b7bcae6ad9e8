"""ConsistencyIndex (MyML/MyML/validation/consistency.py:20-130).

Greedy cluster matching: repeatedly take the pair (i, j) of remaining
clusters with the largest Jaccard coefficient A / (B + C - A) (A the shared
samples, B and C the cluster sizes; float32 arithmetic as in the reference,
which converts its binary matrices to float32, :62), add its A to the shared
count and remove both clusters, min(K1, K2) times; accuracy = shared / N.

The coefficients never change as clusters are removed, so the reference's
rescans (:87-109: strict >, first maximum in row-major order of the remaining
clusters) equal one sweep over all pairs sorted by (coefficient descending,
i, j) that takes a pair when neither cluster is taken yet. When no remaining
pair has a positive coefficient the reference's best pair stays (-1, -1): it
adds -1 and removes the last remaining cluster of each partition (:111-119);
that is kept. A, B and C come from the contingency table counted on the GPU.

The reference's score() passes N= to converters that take no such argument
(MyML/MyML/utils/partition.py:17, 37), so it raises TypeError as written;
this class implements the documented behaviour, with N defaulting to the
number of samples.
"""
from __future__ import annotations

import numpy as np

from .contingency import contingency


def greedy_match_count(T, size1, size2):
    """The reference's _match_ (consistency.py:70-125) on a contingency table:
    returns n_shared with the reference's types (float32 once a pair is taken)."""
    T = np.asarray(T)
    k1, k2 = T.shape
    A = T.astype(np.float32)
    B = np.asarray(size1).astype(np.float32)[:, None]
    C = np.asarray(size2).astype(np.float32)[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        coef = (A / ((B + C) - A)).astype(np.float32)
    pos = np.nonzero(coef > 0)  # NaN and non-positive coefficients are never taken
    order = np.lexsort((pos[1], pos[0], -coef[pos].astype(np.float64)))
    used1 = np.zeros(k1, bool)
    used2 = np.zeros(k2, bool)
    n_shared = 0
    taken = 0
    steps = min(k1, k2)
    for e in order:
        if taken == steps:
            break
        i, j = int(pos[0][e]), int(pos[1][e])
        if used1[i] or used2[j]:
            continue
        used1[i] = used2[j] = True
        n_shared += A[i, j]
        taken += 1
    for _ in range(steps - taken):  # no positive pair left: best stays (-1, -1)
        n_shared += -1
    return n_shared


class ConsistencyIndex:

    def __init__(self, N=None):
        self.N = N

    def score(self, clusts1, clusts2, format='array', format1=None, format2=None,
              array_has_zero=False, N=None):
        """clusts1, clusts2: the two partitions; format 'array' (labels) or
        'list' (index arrays per cluster); format1/format2 override it."""
        if format1 is None:
            format1 = format
        if format2 is None:
            format2 = format
        for f in (format1, format2):
            if f not in ("array", "list"):
                raise Exception("Format not accepted: {}".format(f))
        T, s1, s2, n = contingency(clusts1, clusts2, format1, format2, self.N)
        if self.N is None:
            self.N = n
        self.contingency_ = T
        self.match_count = greedy_match_count(T, s1, s2)
        self.unmatch_count = self.N - self.match_count
        self.accuracy = float(self.match_count) / self.N
        return self.accuracy
