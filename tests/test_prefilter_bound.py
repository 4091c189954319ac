"""The bound behind the pair kernel's exact row filter (indexgen_amd/csrc/
igc_prefilter.hpp): for equal-length strings A, B and any split point h,

    |p(A) - p(B)|_1 + |s(A) - s(B)|_1 <= 2 * Levenshtein(A, B)

with p, s the symbol counts of A[:h], A[h:].  Checked against the oracle's
distance exhaustively for n <= 5 (every pair, every split point), on random and
near pairs up to n = 32, and shown tight (equality is reached).  A filter
built on a wrong bound would drop edges; the GPU tests compare the filtered
graphs with the oracle as well."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def counts(S, a, b):
    return np.stack([(S[:, a:b] == c).sum(1) for c in range(4)], 1)


def split_l1(A, B, h):
    n = A.shape[1]
    return np.abs(counts(A, 0, h) - counts(B, 0, h)).sum(1) + np.abs(counts(A, h, n) - counts(B, h, n)).sum(1)


def distances(A, B):
    return np.array([O.levenshtein(a, b) for a, b in zip(A, B)])


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_bound_exhaustive(n):
    S = np.array(list(itertools.product(range(4), repeat=n)), np.uint8)
    A = np.repeat(S, len(S), axis=0)
    B = np.tile(S, (len(S), 1))
    D = distances(A, B)
    tight = False
    for h in range(0, n + 1):
        L = split_l1(A, B, h)
        assert (L <= 2 * D).all(), (n, h)
        tight |= bool((L == 2 * D)[D > 0].any())
    assert tight or n == 0


def _near(rng, a, edits):
    b = list(a)
    n = len(a)
    for _ in range(edits):
        op = int(rng.integers(0, 3))
        p = int(rng.integers(0, n))
        if op == 0:
            b[p] = int(rng.integers(0, 4))
        elif op == 1:
            del b[p]
            b.append(int(rng.integers(0, 4)))
        else:
            b.insert(p, int(rng.integers(0, 4)))
            b.pop()
    return b


@pytest.mark.parametrize("n", [8, 13, 16, 24, 32])
def test_bound_random_and_near(n):
    rng = np.random.default_rng(70 + n)
    M = 6000
    A = rng.integers(0, 4, size=(M, n), dtype=np.uint8)
    B = rng.integers(0, 4, size=(M, n), dtype=np.uint8)
    for r in range(0, M, 2):
        B[r] = _near(rng, A[r], int(rng.integers(1, 10)))
    D = distances(A, B)
    for h in range(0, n + 1):
        assert (split_l1(A, B, h) <= 2 * D).all(), (n, h)


def test_bound_cannot_be_halved():
    """A one-symbol rotation: distance 2, split distance 4 at the middle."""
    a = np.array([[0, 0, 0, 1, 1, 1]], np.uint8)
    b = np.array([[0, 0, 1, 1, 1, 0]], np.uint8)  # delete a 0 at the front, append one
    assert O.levenshtein(a[0], b[0]) == 2
    assert split_l1(a, b, 3)[0] == 4
