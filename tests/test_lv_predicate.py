"""The kernels' pair predicate (indexgen_amd/csrc/igc_lv.hpp), compiled for the
host with g++, against the oracle: exhaustive over all pairs of short strings
and every threshold, and on random and near pairs up to n = 16."""
import ctypes
import itertools
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lv(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("lv") / "lv_harness.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "lv_harness.cpp")],
                   check=True)
    L = ctypes.CDLL(so)
    L.lv_pairs.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.c_void_p]
    L.lv_pairs.restype = ctypes.c_int

    def pairs(A, B, K):
        A = np.ascontiguousarray(A, np.uint8)
        B = np.ascontiguousarray(B, np.uint8)
        out = np.zeros(A.shape[0], np.uint8)
        assert L.lv_pairs(A.ctypes.data, B.ctypes.data, A.shape[0], A.shape[1], K, out.ctypes.data) == 0
        return out

    return pairs


def all_strings(n):
    return np.array(list(itertools.product(range(4), repeat=n)), np.uint8)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_exhaustive(lv, n):
    S = all_strings(n)
    N = S.shape[0]
    A = np.repeat(S, N, axis=0)
    B = np.tile(S, (N, 1))
    for K in range(0, min(n + 2, 16)):
        want = O.conflict_pairs(A, B, K + 1)
        assert np.array_equal(lv(A, B, K), want), (n, K)
    assert lv(A, B, 16).all()  # the every-pair variant


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


@pytest.mark.parametrize("n", [6, 7, 8, 10, 11, 12, 13, 14, 15, 16])
def test_random_and_near_pairs(lv, n):
    rng = np.random.default_rng(100 + n)
    M = 20000
    A = rng.integers(0, 4, size=(M, n), dtype=np.uint8)
    B = rng.integers(0, 4, size=(M, n), dtype=np.uint8)
    for r in range(0, M, 2):
        B[r] = _near(rng, A[r], int(rng.integers(1, 8)))
    for K in range(0, 16):
        want = O.conflict_pairs(A, B, K + 1)
        got = lv(A, B, K)
        assert np.array_equal(got, want), (n, K, int((got != want).sum()))
