"""HungarianIndex (MyML/MyML/validation/hindex.py:8-86).

Optimal one-to-one cluster matching: the match matrix (shared samples of
every cluster pair, counted on the GPU as a contingency table) is turned into
costs max + 1 - M (:53) and assigned (the reference pads it to square first,
pad_to_square :69-86, which cannot change the optimum); the accuracy is the
matched total over N. The reference assigns with the third-party `munkres`
package (not installed here); any optimal assignment gives the same matched
total, so scipy's linear_sum_assignment replaces it.

Differences from the reference as written: its score() passes N= to
converters that take no such argument (TypeError for 'array' and 'list'),
and its binary matrices are uint8, so its match matrix wraps modulo 256;
here the counts are exact.
"""
from __future__ import annotations

import numpy as np

from .contingency import contingency


def pad_to_square(mat):
    """hindex.py:69-86."""
    if mat.ndim != 2:
        raise TypeError("Array must have two dimensions")
    nrows, ncols = mat.shape
    dif = nrows - ncols
    if dif == 0:
        return mat
    if dif > 0:
        return np.pad(mat, ((0, 0), (0, dif)), mode='constant', constant_values=0)
    return np.pad(mat, ((0, -dif), (0, 0)), mode='constant', constant_values=0)


def hungarian_match_count(match_matrix):
    """Matched total of an optimal assignment. The counts are >= 0, so the
    zero-padded square problem and the rectangular one (min(K1, K2) pairs)
    have the same optimum; the rectangular one is solved (O(K1 K2 min(K1, K2))
    instead of O(max(K1, K2)^3))."""
    from scipy.optimize import linear_sum_assignment
    M = np.asarray(match_matrix, dtype=np.int64)
    if M.size == 0:
        return 0
    profit = M.max() + 1 - M  # hindex.py:53
    rows, cols = linear_sum_assignment(profit)
    return int(M[rows, cols].sum())


class HungarianIndex():

    def __init__(self, nsamples):
        self.N = nsamples

    def score(self, clusts1, clusts2, format='array', format1=None, format2=None):
        if format1 is None:
            format1 = format
        if format2 is None:
            format2 = format
        for f in (format1, format2):
            if f not in ("array", "list", "bin"):
                raise ValueError("Format not accepted: {}".format(f))
        T, _, _, n = contingency(clusts1, clusts2, format1, format2, self.N)
        if self.N is None:
            self.N = n
        self.match_matrix_ = T
        self.match_count = hungarian_match_count(T)
        self.accuracy = float(self.match_count) / self.N
        return self.accuracy
