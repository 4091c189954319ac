"""ctypes front end of the C oracle (oracle/igc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package
(indexgen_amd) never imports this module.

Every function takes symbol matrices as numpy uint8 arrays of shape (N, n)
with values 0..3 and mirrors a reference routine (see the C file's header for
the file:line each one restates).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_DRIVER = os.path.join(_HERE, "_ref", "ref_driver")

_lib = None


def build() -> None:
    """Compile liboracle.so (and _ref/ref_driver where the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle"], check=True)
    subprocess.run(["make", "-s", "-C", _HERE, "-j8", "ref"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            subprocess.run(["make", "-s", "-C", _HERE, "liboracle"], check=True)
        L = ctypes.CDLL(_LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.orc_levenshtein.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int]
        L.orc_levenshtein.restype = ctypes.c_int
        L.orc_banded.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int]
        L.orc_banded.restype = ctypes.c_int
        L.orc_conflict.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int]
        L.orc_conflict.restype = ctypes.c_int
        L.orc_edge_count.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_edge_count.restype = ctypes.c_int64
        L.orc_graph_csr.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p]
        L.orc_graph_csr.restype = ctypes.c_int64
        L.orc_greedy.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_greedy.restype = ctypes.c_int64
        L.orc_max_run.argtypes = [u8p, ctypes.c_int]
        L.orc_max_run.restype = ctypes.c_int
        L.orc_gc_ok.argtypes = [u8p, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_gc_ok.restype = ctypes.c_int
        L.orc_filter.argtypes = [u8p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_void_p]
        L.orc_filter.restype = ctypes.c_int64
        L.orc_check_exhaustive.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_check_exhaustive.restype = ctypes.c_int64
        L.orc_row_neighbors.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_row_neighbors.restype = ctypes.c_int64
        L.orc_conflict_matrix.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_conflict_matrix.restype = None
        L.orc_conflict_pairs.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_conflict_pairs.restype = None
        L.orc_kmeans_seeds.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
        L.orc_kmeans_seeds.restype = None
        L.orc_kmeans.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                 ctypes.c_void_p, ctypes.c_void_p]
        L.orc_kmeans.restype = None
        L.orc_assign_fp32.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.orc_assign_fp32.restype = None
        d, vp, i, i64 = ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        L.orc_gen_linear.argtypes = [vp, i, i, vp, i, d, d, vp, i64]
        L.orc_gen_vt.argtypes = [i, i, i, i, d, d, vp, i64]
        L.orc_gen_dvt.argtypes = [i, i, i, d, d, vp, i64]
        L.orc_gen_all.argtypes = [i, i, d, d, vp, i64]
        for f in (L.orc_gen_linear, L.orc_gen_vt, L.orc_gen_dvt, L.orc_gen_all):
            f.restype = ctypes.c_int64
        _lib = L
    return _lib


def _u8(a: np.ndarray):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def levenshtein(a, b) -> int:
    a, pa = _u8(np.asarray(a))
    b, pb = _u8(np.asarray(b))
    return lib().orc_levenshtein(pa, a.size, pb, b.size)


def banded(text, pattern, k: int) -> int:
    """EditDistanceBanded(text, MakePattern(pattern), k)."""
    t, pt = _u8(np.asarray(text))
    p, pp = _u8(np.asarray(pattern))
    return lib().orc_banded(pt, t.size, pp, p.size, int(k))


def conflict(a, b, min_ed: int) -> bool:
    a, pa = _u8(np.asarray(a))
    b, pb = _u8(np.asarray(b))
    assert a.size == b.size
    return bool(lib().orc_conflict(pa, pb, a.size, int(min_ed)))


def edge_count(syms: np.ndarray, min_ed: int, workers: int = 8) -> int:
    s, ps = _u8(syms)
    N, n = s.shape
    return int(lib().orc_edge_count(ps, N, n, int(min_ed), int(workers)))


def graph_csr(syms: np.ndarray, min_ed: int, T: int = 1, workers: int = 8):
    """Conflict graph as (row_ptr uint64[N+1], col uint32[2|E|]) in the
    neighbour order of the reference's T-thread FillAdjList merge."""
    s, ps = _u8(syms)
    N, n = s.shape
    row_ptr = np.zeros(N + 1, dtype=np.uint64)
    e = lib().orc_graph_csr(ps, N, n, int(min_ed), int(T), int(workers), row_ptr.ctypes.data, None)
    col = np.zeros(max(int(row_ptr[-1]), 1), dtype=np.uint32)
    e2 = lib().orc_graph_csr(ps, N, n, int(min_ed), int(T), int(workers), row_ptr.ctypes.data, col.ctypes.data)
    assert e == e2
    return row_ptr, col[: int(row_ptr[-1])]


def greedy(row_ptr: np.ndarray, col: np.ndarray) -> np.ndarray:
    """Selected node ids in selection order (AdjList greedy + Codebook)."""
    N = row_ptr.size - 1
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    c = np.ascontiguousarray(col, dtype=np.uint32)
    if c.size == 0:
        c = np.zeros(1, dtype=np.uint32)
    order = np.zeros(max(N, 1), dtype=np.int32)
    k = lib().orc_greedy(N, rp.ctypes.data, c.ctypes.data, order.ctypes.data)
    return order[:k].copy()


def max_run(s) -> int:
    s, ps = _u8(np.asarray(s))
    return lib().orc_max_run(ps, s.size)


def gc_ok(s, min_gc: float, max_gc: float) -> bool:
    s, ps = _u8(np.asarray(s))
    return bool(lib().orc_gc_ok(ps, s.size, float(min_gc), float(max_gc)))


def filter_indices(syms: np.ndarray, max_run_: int, min_gc: float, max_gc: float) -> np.ndarray:
    s, ps = _u8(syms)
    N, n = s.shape
    keep = np.zeros(max(N, 1), dtype=np.int64)
    k = lib().orc_filter(ps, N, n, int(max_run_), float(min_gc), float(max_gc), keep.ctypes.data)
    return keep[:k].copy()


def check_exhaustive(n: int, kmax: int) -> int:
    """Mismatches of the banded restatement against min(Levenshtein, k+1) over
    all pairs of length-n strings and k <= kmax (0 expected)."""
    return int(lib().orc_check_exhaustive(int(n), int(kmax)))


def row_neighbors(syms: np.ndarray, min_ed: int, i: int) -> np.ndarray:
    s, ps = _u8(syms)
    N, n = s.shape
    out = np.zeros(max(N, 1), dtype=np.uint32)
    c = lib().orc_row_neighbors(ps, N, n, int(min_ed), int(i), out.ctypes.data)
    return out[:c].copy()


def conflict_matrix(syms: np.ndarray, min_ed: int) -> np.ndarray:
    s, ps = _u8(syms)
    N, n = s.shape
    out = np.zeros((N, N), dtype=np.uint8)
    lib().orc_conflict_matrix(ps, N, n, int(min_ed), out.ctypes.data)
    return out


def conflict_pairs(A: np.ndarray, B: np.ndarray, min_ed: int) -> np.ndarray:
    a, pa = _u8(A)
    b, pb = _u8(B)
    assert a.shape == b.shape
    out = np.zeros(a.shape[0], dtype=np.uint8)
    lib().orc_conflict_pairs(pa, pb, a.shape[0], a.shape[1], int(min_ed), out.ctypes.data)
    return out


def kmeans_seeds(N: int, k: int, seed: int) -> np.ndarray:
    out = np.zeros(k, dtype=np.uint32)
    lib().orc_kmeans_seeds(int(N), int(k), int(seed), out.ctypes.data)
    return out


def kmeans(syms: np.ndarray, k: int, iters: int, seed: int):
    """(assign int32[N], centroids float32[k, 4n]) with the GPU kernels' arithmetic."""
    s, ps = _u8(syms)
    N, n = s.shape
    assign = np.zeros(N, dtype=np.int32)
    C = np.zeros((k, 4 * n), dtype=np.float32)
    lib().orc_kmeans(ps, N, n, int(k), int(iters), int(seed), assign.ctypes.data, C.ctypes.data)
    return assign, C


def assign_fp32(syms: np.ndarray, C: np.ndarray) -> np.ndarray:
    """Nearest-centroid ids (int32[N]) for FP32 centroids C[k, 4n], first index on ties."""
    s, ps = _u8(syms)
    N, n = s.shape
    C = np.ascontiguousarray(C, dtype=np.float32)
    assert C.shape[1] == 4 * n
    out = np.zeros(N, dtype=np.int32)
    lib().orc_assign_fp32(ps, N, n, C.ctypes.data, C.shape[0], out.ctypes.data)
    return out


def _gen(call, n):
    """Runs an orc_gen_* restatement twice: once to count, once to fill."""
    cnt = call(None, 0)
    if cnt < 0:
        raise ValueError("generator arguments out of range")
    out = np.zeros((max(cnt, 1), n), dtype=np.uint8)
    assert call(out.ctypes.data, cnt) == cnt
    return out[:cnt]


def gen_linear(G: np.ndarray, bias=None, max_run: int = 0, min_gc: float = 0.0, max_gc: float = 0.0) -> np.ndarray:
    """Candidates of the LinearCode generator for generator matrix G (k x n over
    GF(4)) and bias, after the filters, in the reference's order."""
    G = np.ascontiguousarray(G, dtype=np.uint8)
    k, n = G.shape
    b = None if bias is None else np.ascontiguousarray(bias, dtype=np.uint8)
    bp = None if b is None else b.ctypes.data
    return _gen(lambda o, c: lib().orc_gen_linear(G.ctypes.data, k, n, bp, max_run, min_gc, max_gc, o, c), n)


def gen_vt(n: int, a: int, b: int, max_run: int = 0, min_gc: float = 0.0, max_gc: float = 0.0) -> np.ndarray:
    return _gen(lambda o, c: lib().orc_gen_vt(n, a, b, max_run, min_gc, max_gc, o, c), n)


def gen_dvt(n: int, s: int, max_run: int = 0, min_gc: float = 0.0, max_gc: float = 0.0) -> np.ndarray:
    return _gen(lambda o, c: lib().orc_gen_dvt(n, s, max_run, min_gc, max_gc, o, c), n)


def gen_all(n: int, max_run: int = 0, min_gc: float = 0.0, max_gc: float = 0.0) -> np.ndarray:
    return _gen(lambda o, c: lib().orc_gen_all(n, max_run, min_gc, max_gc, o, c), n)
