"""Accuracy indices: host matching logic against the literal restatement of
the reference (CPU only; the contingency table is built with numpy here)."""
import numpy as np
import pytest

from masters_final_amd.validation.consistency import greedy_match_count
from masters_final_amd.validation.hindex import hungarian_match_count, pad_to_square
from oracle.validation_oracle import consistency_literal, hungarian_bruteforce


def _table(a, b):
    a, b = np.asarray(a), np.asarray(b)
    T = np.zeros((a.max() - a.min() + 1, b.max() - b.min() + 1), np.int64)
    np.add.at(T, (a - a.min(), b - b.min()), 1)
    return T


def _gapless(x):
    return np.unique(x, return_inverse=True)[1]


@pytest.mark.parametrize("seed", range(6))
def test_greedy_equals_literal_reference_loop(seed):
    rng = np.random.default_rng(seed)
    for _ in range(60):
        n = int(rng.integers(4, 50))
        a = _gapless(rng.integers(0, int(rng.integers(1, 9)), n))
        b = _gapless(rng.integers(0, int(rng.integers(1, 9)), n))
        if rng.random() < 0.3:
            a = a + 1  # 1-based labels
        T = _table(a, b)
        got = greedy_match_count(T, T.sum(1), T.sum(0))
        want, _ = consistency_literal(a, b, n)
        assert got == want and np.float32(got) == np.float32(want)


def test_greedy_no_positive_pair_left():
    """consistency.py:87-119: with no positive coefficient left the best pair
    stays (-1, -1): -1 is added and the last clusters are removed."""
    a, b = np.array([0, 0, 1, 2]), np.array([0, 1, 2, 2])
    T = _table(a, b)
    assert greedy_match_count(T, T.sum(1), T.sum(0)) == 1.0  # 1 + 1 - 1
    assert consistency_literal(a, b, 4)[0] == 1.0


def test_indices_closed_forms():
    rng = np.random.default_rng(7)
    a = _gapless(rng.integers(0, 12, 500))
    perm = rng.permutation(a.max() + 1)
    T = _table(a, perm[a])  # a relabelling of the same partition
    assert greedy_match_count(T, T.sum(1), T.sum(0)) == 500
    assert hungarian_match_count(T) == 500
    b = _gapless(rng.integers(0, 9, 500))
    T = _table(a, b)
    # the optimum is at least the greedy matching and at most N
    assert greedy_match_count(T, T.sum(1), T.sum(0)) <= hungarian_match_count(T) <= 500


@pytest.mark.parametrize("seed", range(4))
def test_hungarian_equals_bruteforce(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(40):
        M = rng.integers(0, 30, (int(rng.integers(1, 7)), int(rng.integers(1, 7))))
        assert hungarian_match_count(M) == hungarian_bruteforce(M)


def test_pad_to_square():
    assert pad_to_square(np.ones((2, 3))).shape == (3, 3)
    assert pad_to_square(np.ones((4, 1))).shape == (4, 4)
    with pytest.raises(TypeError):
        pad_to_square(np.ones(3))
