"""Pins the CPU oracle (oracle/igc_oracle.c) before anything is compared with it.

- banded predicate: against the compiled reference's EditDistanceBanded on
  golden vectors, and exhaustively against Wagner-Fischer for short strings;
- Wagner-Fischer: against the reference's FastEditDistance values and closed
  forms that a dropped term or a swapped index would break;
- graph builder + greedy: against the reference's own FillAdjList adjacency
  (list order) and Codebook selection order for T = 1 and T = 4;
- filters: against direct definitions, including the double-precision
  boundary fractions.
"""
import functools
import os

import numpy as np
import pytest

from oracle import oracle as O


def test_banded_matches_reference_golden(golden_dir):
    z = dict(np.load(os.path.join(golden_dir, "pairs_golden.npz")))  # decompress once
    bad = 0
    for r in range(z["k"].size):
        t = z["text"][r, : z["text_len"][r]]
        p = z["pat"][r, : z["pat_len"][r]]
        if O.banded(t, p, int(z["k"][r])) != int(z["banded"][r]):
            bad += 1
        if O.levenshtein(t, p) != int(z["wf"][r]):
            bad += 1
    assert bad == 0
    # the golden set must exercise both outcomes around every threshold
    assert (z["banded"] <= z["k"]).sum() > 1000 and (z["banded"] > z["k"]).sum() > 1000


@pytest.mark.parametrize("n,kmax", [(1, 3), (2, 4), (3, 5), (4, 6), (5, 6)])
def test_banded_exhaustive_equals_levenshtein(n, kmax):
    assert O.check_exhaustive(n, kmax) == 0


@functools.lru_cache(maxsize=None)
def _lev_rec(a: tuple, b: tuple) -> int:
    # the textbook recurrence, written independently of the DP in C
    if not a:
        return len(b)
    if not b:
        return len(a)
    return min(_lev_rec(a[1:], b) + 1, _lev_rec(a, b[1:]) + 1, _lev_rec(a[1:], b[1:]) + (a[0] != b[0]))


def test_levenshtein_brute_force_small():
    rng = np.random.default_rng(1)
    for _ in range(400):
        la, lb = rng.integers(0, 7, 2)
        a = rng.integers(0, 4, la).astype(np.uint8)
        b = rng.integers(0, 4, lb).astype(np.uint8)
        assert O.levenshtein(a, b) == _lev_rec(tuple(a), tuple(b))


def test_levenshtein_closed_forms():
    rng = np.random.default_rng(2)
    for _ in range(300):
        n = int(rng.integers(1, 17))
        a = rng.integers(0, 4, n).astype(np.uint8)
        b = rng.integers(0, 4, n).astype(np.uint8)
        c = rng.integers(0, 4, n).astype(np.uint8)
        assert O.levenshtein(a, a) == 0
        assert O.levenshtein(a, np.zeros(0, np.uint8)) == n
        dab = O.levenshtein(a, b)
        assert dab == O.levenshtein(b, a)
        assert dab <= int((a != b).sum())  # Hamming upper bound
        cnt = lambda s: np.bincount(s, minlength=4)
        assert 2 * dab >= int(np.abs(cnt(a) - cnt(b)).sum())  # symbol-count lower bound
        assert O.levenshtein(a, c) <= dab + O.levenshtein(b, c)  # triangle inequality
        # one substitution / insertion / deletion costs exactly one
        s = a.copy()
        p = int(rng.integers(0, n))
        s[p] = (s[p] + 1 + int(rng.integers(0, 3))) % 4
        assert O.levenshtein(a, s) == 1
        assert O.levenshtein(a, np.insert(a, p, int(rng.integers(0, 4)))) == 1
        assert O.levenshtein(a, np.delete(a, p)) == 1
        # a rotation by one costs at most two (delete + insert)
        assert O.levenshtein(a, np.roll(a, 1)) <= 2


def test_conflict_is_threshold_of_levenshtein():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        n = int(rng.integers(1, 17))
        a = rng.integers(0, 4, n).astype(np.uint8)
        b = a.copy()
        for _ in range(int(rng.integers(0, 5))):
            b[int(rng.integers(0, n))] = rng.integers(0, 4)
        ed = int(rng.integers(0, 8))
        assert O.conflict(a, b, ed) == (O.levenshtein(a, b) < ed)


def test_graph_and_greedy_match_reference(graphs_golden):
    for s in graphs_golden:
        for T, (rp, col, words) in s["T"].items():
            orp, ocol = O.graph_csr(s["syms"], s["min_ed"], T=T, workers=4)
            assert np.array_equal(orp, rp)
            assert np.array_equal(ocol, col)  # same neighbour order as AdjList::ToFile
            assert np.array_equal(s["syms"][O.greedy(orp, ocol)], words)


def _independent_and_maximal(rp, col, order, N):
    sel = np.zeros(N, bool)
    sel[order] = True
    for v in order:
        assert not sel[col[rp[v]:rp[v + 1]]].any()
    for v in np.flatnonzero(~sel):
        assert sel[col[rp[v]:rp[v + 1]]].any()


def test_greedy_invariants_random_graphs():
    rng = np.random.default_rng(4)
    for _ in range(5):
        N = int(rng.integers(50, 400))
        syms = rng.integers(0, 4, size=(N, 7), dtype=np.uint8)
        rp, col = O.graph_csr(syms, 3)
        order = O.greedy(rp, col)
        assert len(set(order.tolist())) == order.size
        _independent_and_maximal(rp, col, order, N)


def test_published_edge_count_small(golden_dir):
    z = np.load(os.path.join(golden_dir, "published_N11-ED5-GC0_3-0_7-RUN0.npz"))
    rp, col = O.graph_csr(z["syms"], int(z["min_ed"]))
    assert col.size == int(z["ones"])


def _max_run_py(s):
    best = run = 1
    for i in range(1, len(s)):
        run = run + 1 if s[i] == s[i - 1] else 1
        best = max(best, run)
    return best


def test_filters():
    rng = np.random.default_rng(5)
    for _ in range(500):
        n = int(rng.integers(1, 17))
        s = rng.integers(0, 4, n).astype(np.uint8)
        assert O.max_run(s) == _max_run_py(list(s))
        gc = int(((s == 1) | (s == 2)).sum())
        assert O.gc_ok(s, 0.3, 0.7) == (gc / n >= 0.3 and gc / n <= 0.7)
    # boundary fractions in double precision: 3/10 == 0.3 and 7/10 == 0.7 pass
    s = np.array([1, 2, 1, 0, 0, 0, 3, 3, 0, 3], np.uint8)
    assert O.gc_ok(s, 0.3, 0.7)
    s = np.array([1, 2, 1, 2, 1, 2, 1, 0, 0, 3], np.uint8)
    assert O.gc_ok(s, 0.3, 0.7)
    s = np.array([1, 2, 0, 0, 0, 3, 3, 0, 3, 0], np.uint8)
    assert not O.gc_ok(s, 0.3, 0.7)
    syms = rng.integers(0, 4, size=(300, 11), dtype=np.uint8)
    keep = O.filter_indices(syms, 3, 0.3, 0.7)
    want = [i for i in range(300) if _max_run_py(list(syms[i])) <= 3 and O.gc_ok(syms[i], 0.3, 0.7)]
    assert keep.tolist() == want
    assert O.filter_indices(syms, 0, 0.0, 0.0).tolist() == list(range(300))
