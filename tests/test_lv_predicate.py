"""The kernels' pair predicates (indexgen_amd/csrc/igc_lv.hpp), compiled for the
host with g++, against the oracle: Landau-Vishkin on thermometer codes and the
Myers recurrence, each on two-pairs-per-register 16-bit lanes (n <= 16) and on
32-bit lanes (n <= 32), exhaustive over all pairs of short strings and every
threshold, and on random and near pairs up to n = 32; the binary forms on
64-bit lanes (symbols 0/1, n <= 64) the same way.  Also the dispatch rule:
which predicate runs for (n, K)."""
import ctypes
import itertools
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("lv") / "lv_harness.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "lv_harness.cpp")],
                   check=True)
    L = ctypes.CDLL(so)
    for f in (L.lv_pairs, L.myers_pairs):
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.c_void_p]
        f.restype = ctypes.c_int
    for f in (L.lv_pairs_bin, L.myers_pairs_bin):
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        f.restype = ctypes.c_int
    L.pred_kind_of.argtypes = [ctypes.c_int, ctypes.c_int]
    L.pred_kind_of.restype = ctypes.c_int
    L.lv_launch_max_k.restype = ctypes.c_int
    return L


def _call(fn, A, B, K, wide):
    A = np.ascontiguousarray(A, np.uint8)
    B = np.ascontiguousarray(B, np.uint8)
    out = np.zeros(A.shape[0], np.uint8)
    assert fn(A.ctypes.data, B.ctypes.data, A.shape[0], A.shape[1], K, int(wide), out.ctypes.data) == 0
    return out


@pytest.fixture(scope="module")
def lv(harness):
    return lambda A, B, K, wide=False: _call(harness.lv_pairs, A, B, K, wide)


@pytest.fixture(scope="module")
def myers(harness):
    return lambda A, B, K, wide=False: _call(harness.myers_pairs, A, B, K, wide)


def all_strings(n):
    return np.array(list(itertools.product(range(4), repeat=n)), np.uint8)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_exhaustive(lv, myers, n):
    S = all_strings(n)
    N = S.shape[0]
    A = np.repeat(S, N, axis=0)
    B = np.tile(S, (N, 1))
    for K in range(0, min(n + 2, 16)):
        want = O.conflict_pairs(A, B, K + 1)
        for wide in (False, True):
            assert np.array_equal(lv(A, B, K, wide), want), ("lv", n, K, wide)
            assert np.array_equal(myers(A, B, K, wide), want), ("myers", n, K, wide)


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


def _pairs(n, M=20000, seed=0):
    rng = np.random.default_rng(100 + n + seed)
    A = rng.integers(0, 4, size=(M, n), dtype=np.uint8)
    B = rng.integers(0, 4, size=(M, n), dtype=np.uint8)
    for r in range(0, M, 2):
        B[r] = _near(rng, A[r], int(rng.integers(1, 8 if n <= 16 else 14)))
    # a few strings with long runs of one symbol (all-equal planes)
    A[1:200:10] = 0
    B[3:300:10] = 3
    return A, B


@pytest.mark.parametrize("n", [6, 7, 8, 10, 11, 12, 13, 14, 15, 16])
def test_random_and_near_pairs(lv, myers, n):
    A, B = _pairs(n)
    for K in range(0, 16):
        want = O.conflict_pairs(A, B, K + 1)
        for wide in (False, True):
            got = lv(A, B, K, wide)
            assert np.array_equal(got, want), ("lv", n, K, wide, int((got != want).sum()))
        for wide in (False, True):
            got = myers(A, B, K, wide)
            assert np.array_equal(got, want), ("myers", n, K, wide, int((got != want).sum()))


@pytest.mark.parametrize("n", [17, 18, 20, 23, 24, 27, 31, 32])
def test_random_and_near_pairs_wide(lv, myers, n):
    A, B = _pairs(n, M=12000)
    for K in range(0, n + 1):
        want = O.conflict_pairs(A, B, K + 1)
        got = myers(A, B, K, True)
        assert np.array_equal(got, want), ("myers", n, K, int((got != want).sum()))
        if K <= 15:
            got = lv(A, B, K, True)
            assert np.array_equal(got, want), ("lv", n, K, int((got != want).sum()))


def test_oracle_pins_wide_distances():
    """The oracle's distance at n > 16 against the textbook dynamic program."""
    rng = np.random.default_rng(7)
    for n in (17, 24, 32):
        for _ in range(60):
            a = rng.integers(0, 4, n).astype(np.uint8)
            b = np.array(_near(rng, list(a), int(rng.integers(0, 12))), np.uint8)
            D = np.arange(n + 1)
            for i in range(1, n + 1):
                prev, D = D, np.empty(n + 1, np.int64)
                D[0] = i
                for j in range(1, n + 1):
                    D[j] = min(prev[j] + 1, D[j - 1] + 1, prev[j - 1] + (a[i - 1] != b[j - 1]))
            assert O.levenshtein(a, b) == D[n]


def test_dispatch_rule(harness):
    """pred_kind: every pair when K >= n; Landau-Vishkin only up to the largest
    K the kernels are instantiated for, and for every K up to it at some n;
    Landau-Vishkin's reach in K grows with n within a word class (Myers' cost
    grows with n, Landau-Vishkin's does not); at the benchmark shape (n=13,
    K=3) Landau-Vishkin."""
    kmax = harness.lv_launch_max_k()
    seen, lv_k = set(), set()
    reach = {}
    for n in range(1, 65):
        for K in range(0, 70):
            k = harness.pred_kind_of(n, K)
            seen.add(k)
            if K >= n:
                assert k == 2, (n, K)
                continue
            assert k in (0, 1), (n, K)
            if k == 0:
                assert K <= kmax, (n, K)
                lv_k.add(K)
                # Landau-Vishkin at K implies it at every smaller K
                assert all(harness.pred_kind_of(n, j) == 0 for j in range(K)), (n, K)
                reach[n] = K
    for lo, hi in ((1, 16), (17, 32), (33, 64)):
        r = [reach.get(n, -1) for n in range(lo, hi + 1)]
        assert r == sorted(r), (lo, hi, r)
    assert seen == {0, 1, 2}
    assert lv_k == set(range(kmax + 1))
    assert harness.pred_kind_of(13, 3) == 0
    assert harness.pred_kind_of(32, 3) == 0
    assert harness.pred_kind_of(16, 14) == 1

def _bin_call(fn, A, B, K):
    A = np.ascontiguousarray(A, np.uint8)
    B = np.ascontiguousarray(B, np.uint8)
    out = np.zeros(A.shape[0], np.uint8)
    assert fn(A.ctypes.data, B.ctypes.data, A.shape[0], A.shape[1], K, out.ctypes.data) == 0
    return out


def _bin_near(rng, a, edits):
    b = list(a)
    n = len(a)
    for _ in range(edits):
        op = int(rng.integers(0, 3))
        p = int(rng.integers(0, n))
        if op == 0:
            b[p] ^= 1
        elif op == 1:
            del b[p]
            b.append(int(rng.integers(0, 2)))
        else:
            b.insert(p, int(rng.integers(0, 2)))
            b.pop()
    return b


def _bin_pairs(n, M, seed=0):
    rng = np.random.default_rng(500 + n + seed)
    A = rng.integers(0, 2, size=(M, n), dtype=np.uint8)
    B = rng.integers(0, 2, size=(M, n), dtype=np.uint8)
    for r in range(0, M, 2):
        B[r] = _bin_near(rng, A[r], int(rng.integers(1, 18)))
    A[1:200:10] = 0
    B[3:300:10] = 1
    A[5:400:20] = 1
    return A, B


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_binary_exhaustive(harness, n):
    S = np.array(list(itertools.product(range(2), repeat=n)), np.uint8)
    N = S.shape[0]
    A = np.repeat(S, N, axis=0)
    B = np.tile(S, (N, 1))
    for K in range(0, min(n + 2, 16)):
        want = O.conflict_pairs(A, B, K + 1)
        assert np.array_equal(_bin_call(harness.lv_pairs_bin, A, B, K), want), ("lv", n, K)
        assert np.array_equal(_bin_call(harness.myers_pairs_bin, A, B, K), want), ("myers", n, K)


@pytest.mark.parametrize("n", [17, 31, 32, 33, 34, 40, 47, 48, 57, 63, 64])
def test_binary_random_and_near_pairs(harness, n):
    A, B = _bin_pairs(n, 6000)
    for K in range(0, n + 1):
        want = O.conflict_pairs(A, B, K + 1)
        got = _bin_call(harness.myers_pairs_bin, A, B, K)
        assert np.array_equal(got, want), ("myers", n, K, int((got != want).sum()))
        if K <= 15:
            got = _bin_call(harness.lv_pairs_bin, A, B, K)
            assert np.array_equal(got, want), ("lv", n, K, int((got != want).sum()))


def test_oracle_pins_binary_distances():
    """The oracle's distance and banded predicate on binary strings up to
    n = 64 against the textbook dynamic program."""
    rng = np.random.default_rng(11)
    for n in (33, 48, 64):
        for _ in range(40):
            a = rng.integers(0, 2, n).astype(np.uint8)
            b = np.array(_bin_near(rng, list(a), int(rng.integers(0, 20))), np.uint8)
            D = np.arange(n + 1)
            for i in range(1, n + 1):
                prev, D = D, np.empty(n + 1, np.int64)
                D[0] = i
                for j in range(1, n + 1):
                    D[j] = min(prev[j] + 1, D[j - 1] + 1, prev[j - 1] + (a[i - 1] != b[j - 1]))
            assert O.levenshtein(a, b) == D[n]
            for K in (0, 1, 3, int(D[n]) - 1, int(D[n]), int(D[n]) + 1):
                if K < 0:
                    continue
                assert O.conflict(a, b, K + 1) == (D[n] <= K), (n, K, int(D[n]))


def test_binary_dispatch_rule(harness):
    """Binary lengths: every pair once K >= n; Landau-Vishkin at every
    instantiated K from n = 40 on (Myers' 64-bit steps cost n each), and at
    every K up to 12 from n = 33."""
    kmax = harness.lv_launch_max_k()
    for n in range(33, 65):
        for K in range(0, 70):
            k = harness.pred_kind_of(n, K)
            if K >= n:
                assert k == 2
            elif K > kmax:
                assert k == 1, (n, K)
            elif K <= 12 or n >= 40:
                assert k == 0, (n, K)
