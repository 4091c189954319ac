"""Pins the oracle's FP32 K-means (oracle/igc_oracle.c, orc_kmeans) -- which has
no reference counterpart -- against scikit-learn's Lloyd algorithm started from
the same centroids (an independent textbook implementation in float64), and
checks the library's seed selection matches the oracle's."""
import numpy as np
import pytest

import indexgen_amd as igc
from oracle import oracle as O


def one_hot(syms):
    N, n = syms.shape
    X = np.zeros((N, 4 * n), dtype=np.float64)
    X[np.repeat(np.arange(N), n), (4 * np.arange(n))[None, :].repeat(N, 0).ravel() + syms.ravel()] = 1.0
    return X


def lloyd_f64(X, init, iters):
    """Textbook Lloyd in float64 on the explicit feature vectors: squared
    Euclidean distances, first index on ties, mean update, empty clusters keep
    their centroid, final assignment after the last update."""
    C = init.copy()
    for _ in range(iters):
        lab = ((X[:, None, :] - C[None]) ** 2).sum(-1).argmin(1)
        for c in range(C.shape[0]):
            m = lab == c
            if m.any():
                C[c] = X[m].mean(0)
    d = ((X[:, None, :] - C[None]) ** 2).sum(-1)
    return d.argmin(1), C, d


@pytest.mark.parametrize("N,n,k,iters,seed", [(3000, 16, 16, 10, 1), (2500, 11, 8, 6, 2), (4000, 13, 32, 10, 3)])
def test_oracle_kmeans_matches_textbook_lloyd(N, n, k, iters, seed):
    rng = np.random.default_rng(seed)
    syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
    assign, C = O.kmeans(syms, k, iters, seed)
    X = one_hot(syms)
    init = X[O.kmeans_seeds(N, k, seed)]
    lab, C64, d = lloyd_f64(X, init, iters)
    diff = np.flatnonzero(lab != assign)
    # float32 and float64 can only disagree on near-ties
    for i in diff:
        assert abs(d[i, assign[i]] - d[i, lab[i]]) < 1e-5
    assert diff.size <= N // 1000
    assert np.abs(C.astype(np.float64) - C64).max() < 1e-5 or diff.size > 0


def test_oracle_kmeans_inertia_close_to_sklearn():
    from sklearn.cluster import KMeans
    rng = np.random.default_rng(11)
    syms = rng.integers(0, 4, size=(3000, 12), dtype=np.uint8)
    assign, C = O.kmeans(syms, 12, 10, 4)
    X = one_hot(syms)
    km = KMeans(n_clusters=12, init=X[O.kmeans_seeds(3000, 12, 4)], n_init=1, max_iter=10, tol=0.0,
                algorithm="lloyd").fit(X)
    ours = ((X - C[assign].astype(np.float64)) ** 2).sum()
    assert abs(ours - km.inertia_) / km.inertia_ < 0.02


def test_seed_selection_matches_oracle():
    for N, k, seed in [(10, 10, 0), (1000, 37, 5), (500000, 256, 4)]:
        assert np.array_equal(igc.kmeans_seeds(N, k, seed), O.kmeans_seeds(N, k, seed))
        assert np.unique(O.kmeans_seeds(N, k, seed)).size == k


def test_assign_fp32_oracle_matches_brute_force():
    """orc_assign_fp32 (the GPU assignment's arithmetic) against a float64
    brute-force argmin over explicit one-hot vectors, with centroids whose
    entries are multiples of 1/8 (exact in float32, so no rounding ties)."""
    rng = np.random.default_rng(61)
    for N, n, k in [(3000, 13, 37), (1000, 24, 9), (500, 5, 60)]:
        syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        C = (rng.integers(0, 9, size=(k, 4 * n)) / 8.0).astype(np.float32)
        C[: min(k, 5)] = 0.0
        for c in range(min(k, 5)):  # one-hot copies of real strings: exact ties with other rows
            C[c, 4 * np.arange(n) + syms[c]] = 1.0
        X = np.zeros((N, 4 * n))
        X[np.repeat(np.arange(N), n), (4 * np.arange(n))[None, :].repeat(N, 0).ravel() + syms.ravel()] = 1.0
        d = ((X[:, None, :] - C[None, :, :].astype(np.float64)) ** 2).sum(-1)
        want = d.argmin(1)  # first index on ties
        got = O.assign_fp32(syms, C)
        assert np.array_equal(got, want), (N, n, k)
