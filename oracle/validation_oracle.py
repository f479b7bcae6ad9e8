"""Literal restatement of the reference's accuracy indices -- TEST
INFRASTRUCTURE ONLY (tests/ import it as the checker; the product never does).

consistency_literal follows MyML/MyML/validation/consistency.py:70-125 step by
step on the binary float32 cluster-by-sample matrices of
MyML/MyML/utils/partition.py:37-70 (rows in label order, 1-based labels
shifted down), including np.delete of the chosen rows and the (-1, -1) best
pair when no coefficient is positive. Pinned by: identical partitions give N
(closed form), a relabelling gives N, and hand cases.

hungarian_bruteforce is the optimum of MyML/MyML/validation/hindex.py:8-66
(any optimal assignment of the zero-padded match matrix has the same total)
by enumerating every permutation: the definition of the optimum, for k <= 7.
"""
from __future__ import annotations

import itertools

import numpy as np


def to_bin(labels):
    """partition.py:37-70 (array format -> binary cluster-by-sample, float32
    as consistency.py:62 converts it)."""
    lab = np.asarray(labels)
    u = np.unique(lab)
    if u.size not in (lab.max(), lab.max() + 1):
        raise ValueError("labels must be [0, K) or [1, K] without gaps")
    off = -1 if lab.min() == 1 else 0
    out = np.zeros((u.size, lab.size), dtype=np.float32)
    out[lab + off, np.arange(lab.size)] = 1
    return out


def consistency_literal(labels1, labels2, N):
    c1, c2 = to_bin(labels1), to_bin(labels2)
    n1, n2 = c1.shape[0], c2.shape[0]
    n_shared = 0
    for _ in range(min(n1, n2)):
        max_coef = 0
        k, l = -1, -1
        saved = -1
        for i in range(n1):
            for j in range(n2):
                A = c1[i, :].dot(c2[j, :])
                B = c1[i, :].dot(c1[i, :])
                C = c2[j, :].dot(c2[j, :])
                with np.errstate(divide="ignore", invalid="ignore"):
                    coef = A / (B + C - A)
                if coef > max_coef:
                    k, l = i, j
                    saved = A
                    max_coef = coef
        n_shared += saved
        c1 = np.delete(c1, k, axis=0)
        c2 = np.delete(c2, l, axis=0)
        n1 -= 1
        n2 -= 1
    return n_shared, float(n_shared) / N


def hungarian_bruteforce(match_matrix):
    M = np.asarray(match_matrix, dtype=np.int64)
    r, c = M.shape
    n = max(r, c)
    P = np.zeros((n, n), np.int64)
    P[:r, :c] = M
    best = 0
    for perm in itertools.permutations(range(n)):
        best = max(best, int(P[np.arange(n), perm].sum()))
    return best
