"""Seeded synthetic candidate sets for the conflict-graph workloads.

All randomness comes from numpy's PCG64 generator seeded explicitly, so a seed
names a workload exactly.  Filters follow the reference's applyFilters
(src/CandidateGenerator.cpp:24-50): MaxRun (src/Utils.cpp:420-444) and the GC
fraction test in double precision (src/Utils.cpp:447-462).

Workloads (sizes as named by the benchmark configurations):
  plain_small   n=11, editDist=3: 20,000 uniform strings, dedup (first kept),
                maxRun<=3, GC in [0.3,0.7]                       (~14.4k strings)
  linear_large  n=13, editDist=4: codewords of a seeded systematic [13,8]
                GF(4) code with minimum Hamming distance >= 3, filtered, then
                padded with uniform filtered strings to 200,000
  batched       64 sets, n ~ U{10..16}, editDist ~ U{3,4,5},
                sizes ~ clip(lognormal(ln 20000, 1.0), 1000, 150000)
  vt_refine     n=16, editDist=5: 500,000 strings with sum_i i*x_i = 0 mod 17
  serving       Poisson(20/s) stream of 2,000 jobs, n ~ U{11..16},
                editDist ~ U{3,4,5}, sizes ~ clip(lognormal(ln 30000, 1.2), 500, 400000);
                each job's strings from filtered_uniform_keys
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# GF(4) with w a root of x^2+x+1 labelled 0,1,2=w,3=w^2 (tables of src/Utils.cpp:566-578)
GF4_ADD = np.array([[0, 1, 2, 3], [1, 0, 3, 2], [2, 3, 0, 1], [3, 2, 1, 0]], dtype=np.uint8)
GF4_MUL = np.array([[0, 0, 0, 0], [0, 1, 2, 3], [0, 2, 3, 1], [0, 3, 1, 2]], dtype=np.uint8)


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def max_run(syms: np.ndarray) -> np.ndarray:
    """Longest run of equal symbols per row."""
    N, n = syms.shape
    best = np.ones(N, dtype=np.int32)
    run = np.ones(N, dtype=np.int32)
    for p in range(1, n):
        same = syms[:, p] == syms[:, p - 1]
        run = np.where(same, run + 1, 1)
        best = np.maximum(best, run)
    return best


def gc_ok(syms: np.ndarray, min_gc: float, max_gc: float) -> np.ndarray:
    n = syms.shape[1]
    gc = ((syms == 1) | (syms == 2)).sum(axis=1).astype(np.float64)
    frac = gc / np.float64(n)
    return (frac >= min_gc) & (frac <= max_gc)


def filter_mask(syms: np.ndarray, max_run_: int = 3, min_gc: float = 0.3, max_gc: float = 0.7) -> np.ndarray:
    ok = np.ones(syms.shape[0], dtype=bool)
    if max_run_ > 0:
        ok &= max_run(syms) <= max_run_
    if min_gc > 0 or max_gc > 0:
        ok &= gc_ok(syms, min_gc, max_gc)
    return ok


def dedup_first(syms: np.ndarray) -> np.ndarray:
    """Drops repeated strings, keeping each first occurrence, order preserved."""
    if syms.shape[0] == 0:
        return syms
    _, idx = np.unique(syms, axis=0, return_index=True)
    return syms[np.sort(idx)]


def uniform_strings(g: np.random.Generator, count: int, n: int) -> np.ndarray:
    return g.integers(0, 4, size=(count, n), dtype=np.uint8)


def _keys(syms: np.ndarray) -> np.ndarray:
    """One integer per string (base-4 digits), equal iff the strings are equal."""
    n = syms.shape[1]
    w = (4 ** np.arange(n, dtype=np.uint64)).astype(np.uint64)
    return (syms.astype(np.uint64) * w[None, :]).sum(axis=1)


def filtered_uniform(g: np.random.Generator, count: int, n: int, exclude: np.ndarray | None = None,
                     predicate=None) -> np.ndarray:
    """`count` distinct uniform strings passing the filters (and `predicate`),
    none equal to a row of `exclude`, in draw order."""
    chunks, have = [], 0
    seen = np.zeros(0, dtype=np.uint64) if exclude is None else np.unique(_keys(exclude))
    while have < count:
        need = count - have
        batch = uniform_strings(g, max(2 * need, 1024), n)
        ok = filter_mask(batch)
        if predicate is not None:
            ok &= predicate(batch)
        batch = batch[ok]
        k = _keys(batch)
        _, first = np.unique(k, return_index=True)
        first = np.sort(first)            # first occurrence within the draw, in draw order
        batch, k = batch[first], k[first]
        fresh = ~np.isin(k, seen)
        batch, k = batch[fresh][:need], k[fresh][:need]
        if batch.shape[0]:
            chunks.append(batch)
            have += batch.shape[0]
            seen = np.union1d(seen, k)
    return np.concatenate(chunks) if chunks else np.zeros((0, n), dtype=np.uint8)


def _keys_pass(k: np.ndarray, n: int, max_run_: int = 3, min_gc: float = 0.3, max_gc: float = 0.7) -> np.ndarray:
    """Filter of base-4 keys (symbol p in bits 2p, 2p+1): longest run <= max_run_
    and GC fraction in [min_gc, max_gc] (symbols 1 and 2), as filter_mask."""
    nb = 2 * n
    low = np.uint64(int("01" * n, 2))  # bit 2p of each symbol
    ok = np.ones(k.shape[0], dtype=bool)
    if max_run_ > 0 and n > max_run_:
        x = k ^ (k >> np.uint64(2))          # group p is zero iff s_p == s_{p+1}
        same = ~(x | (x >> np.uint64(1))) & np.uint64(int("01" * (n - 1), 2))
        run = same                           # bit 2p: s_p .. s_{p+r} equal
        for r in range(1, max_run_):
            run = run & (same >> np.uint64(2 * r))
        ok &= run == 0
    if min_gc > 0 or max_gc > 0:
        gc = np.bitwise_count((k ^ (k >> np.uint64(1))) & low).astype(np.float64)
        frac = gc / np.float64(n)
        ok &= (frac >= min_gc) & (frac <= max_gc)
    del nb
    return ok


def _keys_to_symbols(k: np.ndarray, n: int) -> np.ndarray:
    sh = (2 * np.arange(n, dtype=np.uint64))[None, :]
    return ((k[:, None] >> sh) & np.uint64(3)).astype(np.uint8)


def filtered_uniform_keys(g: np.random.Generator, count: int, n: int) -> np.ndarray:
    """`count` distinct uniform strings of length n <= 32 passing maxRun <= 3 and
    GC in [0.3, 0.7], in draw order: the same distribution as filtered_uniform,
    drawn as 2n-bit keys and filtered with bit operations (about 20x faster; a
    different sample for the same seed)."""
    assert 1 <= n <= 32
    out, have = [], 0
    seen = np.zeros(0, dtype=np.uint64)
    top = 1 << (2 * n)
    while have < count:
        need = count - have
        k = g.integers(0, top, size=max(2 * need, 1024), dtype=np.uint64) if n < 32 else \
            g.integers(0, 2 ** 63, size=max(2 * need, 1024), dtype=np.uint64) * np.uint64(2) + \
            g.integers(0, 2, size=max(2 * need, 1024), dtype=np.uint64)
        k = k[_keys_pass(k, n)]
        _, first = np.unique(k, return_index=True)
        k = k[np.sort(first)]
        if seen.size:
            k = k[~np.isin(k, seen)]
        k = k[:need]
        if k.size:
            out.append(k)
            have += k.size
            seen = np.union1d(seen, k) if have < count else seen
    keys = np.concatenate(out) if out else np.zeros(0, dtype=np.uint64)
    return _keys_to_symbols(keys, n)


def plain_small(seed: int = 1, count: int = 20000, n: int = 11) -> np.ndarray:
    g = rng(seed)
    s = dedup_first(uniform_strings(g, count, n))
    return s[filter_mask(s)]


def systematic_gf4_code(g: np.random.Generator, n: int = 13, k: int = 8, min_hd: int = 3):
    """Generator matrix [I_k | P] over GF(4), P rejection-sampled until every
    nonzero codeword has Hamming weight >= min_hd; returns (P, codewords)."""
    msgs = np.array(np.unravel_index(np.arange(4 ** k), (4,) * k), dtype=np.uint8).T  # (4^k, k), base-4 digits
    while True:
        P = g.integers(0, 4, size=(k, n - k), dtype=np.uint8)
        par = np.zeros((msgs.shape[0], n - k), dtype=np.uint8)
        for r in range(k):
            par = GF4_ADD[par, GF4_MUL[msgs[:, r:r + 1], P[r][None, :]]]
        words = np.concatenate([msgs, par], axis=1)
        wt = (words != 0).sum(axis=1)
        if wt[1:].min() >= min_hd:
            return P, words


def linear_large(seed: int = 2, total: int = 200000, n: int = 13) -> np.ndarray:
    g = rng(seed)
    _, words = systematic_gf4_code(g, n, 8, 3)
    words = words[filter_mask(words)]
    if words.shape[0] >= total:
        return words[:total].copy()
    pad = filtered_uniform(g, total - words.shape[0], n, exclude=words)
    return np.concatenate([words, pad])


@dataclass
class CandidateSet:
    syms: np.ndarray
    n: int
    edit_dist: int

    @property
    def pairs(self) -> int:
        N = self.syms.shape[0]
        return N * (N - 1) // 2


def batched(seed: int = 3, sets: int = 64, mu: float = np.log(20000), sigma: float = 1.0, lo: int = 1000,
            hi: int = 150000, scale: float = 1.0) -> list:
    """Uneven independent sets; `scale` shrinks every size (for quick tests)."""
    g = rng(seed)
    out = []
    for _ in range(sets):
        n = int(g.integers(10, 17))
        ed = int(g.integers(3, 6))
        size = int(np.clip(np.exp(g.normal(mu, sigma)), lo, hi) * scale)
        size = max(size, 2)
        out.append(CandidateSet(filtered_uniform(g, size, n), n, ed))
    return out


def vt_syndrome_ok(syms: np.ndarray, mod: int = 17) -> np.ndarray:
    w = np.arange(1, syms.shape[1] + 1, dtype=np.int64)
    return (syms.astype(np.int64) @ w) % mod == 0


def vt_refine(seed: int = 4, total: int = 500000, n: int = 16) -> np.ndarray:
    g = rng(seed)
    return filtered_uniform(g, total, n, predicate=vt_syndrome_ok)


@dataclass
class JobSpec:
    arrival_s: float
    n: int
    edit_dist: int
    size: int
    seed: int

    @property
    def pairs(self) -> int:
        return self.size * (self.size - 1) // 2

    def draw(self) -> np.ndarray:
        """The job's candidates: `size` distinct uniform filtered strings."""
        return filtered_uniform_keys(rng(self.seed), self.size, self.n)


def serving(seed: int = 5, jobs: int = 2000, rate: float = 20.0, mu: float = np.log(30000), sigma: float = 1.2,
            lo: int = 500, hi: int = 400000) -> list:
    """Poisson(rate) arrivals of codebook jobs; strings are drawn per job with
    JobSpec.draw() so a long stream need not be held in memory."""
    g = rng(seed)
    t = 0.0
    out = []
    for _ in range(jobs):
        t += float(g.exponential(1.0 / rate))
        n = int(g.integers(11, 17))
        ed = int(g.integers(3, 6))
        size = int(np.clip(np.exp(g.normal(mu, sigma)), lo, hi))
        out.append(JobSpec(t, n, ed, size, int(g.integers(0, 2 ** 62))))
    return out
