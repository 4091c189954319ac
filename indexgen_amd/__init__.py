"""indexgen_amd -- MI355X conflict-graph engine for IndexGen codebook generation.

Python host mirror of the reference's hot path.  The names follow the
reference's C++ host (src/SparseMat.cpp) so a caller of the reference finds the
same steps:

    FillAdjListGPU(adj, candidates, minED, ones&)   -> fill_adj_list_gpu(candidates, min_ed)
    Codebook(adj, codebook&, candidates, ...)       -> codebook(adj, candidates)
    SolveIndependentSet(candidates, minED, ...)     -> solve_independent_set(candidates, min_ed)
    VerifyDist(codebook, minED, ...)                -> verify_dist(codebook, min_ed)
    ToFile(codebook, params, ...)                   -> write_codebook(...)

All graph work runs in libigc.so (hand-written HIP kernels for gfx950, behind
the C ABI of include/igc.h).  There is no CPU fallback: without a gfx950 device
every graph call raises IGCError.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libigc.so")

IGC_OK = 0
IGC_ERR_INVALID = 1
IGC_ERR_OOM = 2
IGC_ERR_DEVICE = 3
IGC_ERR_ADMISSION = 4
IGC_ERR_IO = 5
IGC_ERR_BUSY = 6
IGC_MAX_LEN = 64  # strings longer than IGC_MAX_QUATERNARY_LEN must be binary (symbols 0/1)
IGC_MAX_QUATERNARY_LEN = 32
IGC_UPPER_TRIANGLE = 1  # flag: CSR rows list only neighbours above the row
IGC_RESULT_ON_DEVICE = 2  # flag: igc_graph_build leaves the CSR in HBM
IGC_PREFILTER_OFF, IGC_PREFILTER_AUTO, IGC_PREFILTER_ALWAYS = 0, 1, 2
IGC_PREFILTER_MIN_N = 16384  # smallest set the automatic mode sorts
IGC_CSR_AUTO, IGC_CSR_WIDE_KEYS = 0, 1
IGC_ON_HOST = 0
IGC_ON_DEVICE = 1


def order_threads(T: int) -> int:
    """Flag reproducing the neighbour order of the reference's T-thread FillAdjList."""
    return (int(T) & 0xFFFF) << 16


class IGCError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"igc error {code}: {msg}")
        self.code = code


class _Csr(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_uint32),
        ("location", ctypes.c_uint32),
        ("n_edges", ctypes.c_uint64),
        ("row_ptr", ctypes.POINTER(ctypes.c_uint64)),
        ("col", ctypes.POINTER(ctypes.c_uint32)),
    ]


class _Timing(ctypes.Structure):
    _fields_ = [
        ("pack_ms", ctypes.c_float),
        ("pairs_ms", ctypes.c_float),
        ("csr_ms", ctypes.c_float),
        ("total_ms", ctypes.c_float),
        ("pairs", ctypes.c_uint64),
        ("n_edges", ctypes.c_uint64),
        ("launches", ctypes.c_uint32),
        ("retries", ctypes.c_uint32),
    ]


class _Set(ctypes.Structure):
    _fields_ = [
        ("syms", ctypes.c_void_p),
        ("N", ctypes.c_uint32),
        ("n", ctypes.c_uint32),
        ("edit_dist", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
    ]


IGC_REFINE_MAX_ROUNDS = 32
IGC_KMEANS_EAGER = 1


class _RefineRound(ctypes.Structure):
    _fields_ = [
        ("candidates", ctypes.c_uint32),
        ("clusters", ctypes.c_uint32),
        ("codewords", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("pairs", ctypes.c_uint64),
        ("kmeans_ms", ctypes.c_float),
        ("pairs_ms", ctypes.c_float),
        ("cluster_s", ctypes.c_double),
        ("graph_s", ctypes.c_double),
        ("greedy_s", ctypes.c_double),
    ]


class _RefineStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_uint32), ("converged", ctypes.c_uint32),
                ("round", _RefineRound * IGC_REFINE_MAX_ROUNDS)]


class _JobResult(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("status", ctypes.c_int),
        ("n_nodes", ctypes.c_uint32),
        ("codewords", ctypes.c_uint32),
        ("batch", ctypes.c_uint32),
        ("batch_jobs", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("order", ctypes.POINTER(ctypes.c_uint32)),
        ("n_edges", ctypes.c_uint64),
        ("submit_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("gpu_done_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("greedy_start_s", ctypes.c_double),
    ]


class _ServerInfo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("batches", "jobs_done", "pairs_done", "queued", "in_use_bytes",
                                               "max_in_use_bytes", "max_batch_jobs")] + [("gpu_busy_s",
                                                                                          ctypes.c_double)]


class _TuneEntry(ctypes.Structure):
    _fields_ = [("cpl", ctypes.c_int), ("q", ctypes.c_uint32), ("ms", ctypes.c_float), ("gpairs_s", ctypes.c_float)]


class _PairPlan(ctypes.Structure):
    _fields_ = [("predicate", ctypes.c_int), ("k", ctypes.c_int), ("lane_cols", ctypes.c_int), ("slots", ctypes.c_int),
                ("q", ctypes.c_uint32), ("tile_cols", ctypes.c_uint32), ("tile_rows", ctypes.c_uint32),
                ("tiles", ctypes.c_uint64), ("row_filter", ctypes.c_int), ("pad", ctypes.c_int)]


class _Filter(ctypes.Structure):
    _fields_ = [("max_run", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("min_gc", ctypes.c_double),
                ("max_gc", ctypes.c_double)]


class _Cands(ctypes.Structure):
    _fields_ = [("d_syms", ctypes.c_void_p), ("count", ctypes.c_uint64), ("enumerated", ctypes.c_uint64),
                ("serial", ctypes.c_uint64), ("n", ctypes.c_uint32), ("ms", ctypes.c_float)]


IGC_GEN_MAX_WORD_LEN = 21

PREDICATES = {0: "landau_vishkin", 1: "myers", 2: "all_pairs"}
IGC_PRED_AUTO, IGC_PRED_LANDAU_VISHKIN, IGC_PRED_MYERS = -1, 0, 1


class _CodebookInfo(ctypes.Structure):
    _fields_ = [
        ("code_len", ctypes.c_int),
        ("min_ed", ctypes.c_int),
        ("max_run", ctypes.c_int),
        ("min_gc", ctypes.c_double),
        ("max_gc", ctypes.c_double),
        ("method_line", ctypes.c_char_p),
        ("candidate_num", ctypes.c_int64),
        ("matrix_ones", ctypes.c_int64),
        ("threads", ctypes.c_int),
        ("candidates_s", ctypes.c_double),
        ("fill_s", ctypes.c_double),
        ("process_s", ctypes.c_double),
        ("overall_s", ctypes.c_double),
        ("cluster_k", ctypes.c_int),
        ("cluster_iterations", ctypes.c_int),
        ("convergence_iterations", ctypes.c_int),
        ("cluster_method", ctypes.c_char_p),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    """Loads the in-tree libigc.so (built by indexgen_amd.build / __graft_entry__.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IGCError(IGC_ERR_DEVICE, f"{LIB_PATH} is missing: run `python -m indexgen_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, u8p = ctypes.c_void_p, ctypes.c_void_p
    u32, u64, i = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    sig = {
        "igc_ctx_create": (i, [i, ctypes.POINTER(vp)]),
        "igc_ctx_destroy": (i, [vp]),
        "igc_ctx_set_stream": (i, [vp, vp]),
        "igc_ctx_stream": (vp, [vp]),
        "igc_last_timing": (i, [vp, ctypes.POINTER(_Timing)]),
        "igc_graph_build": (i, [vp, u8p, u32, u32, u32, u32, ctypes.POINTER(_Csr)]),
        "igc_graph_build_device": (i, [vp, u8p, u32, u32, u32, ctypes.POINTER(_Csr)]),
        "igc_graph_build_device_noalloc": (i, [vp, vp, u32, u32, u32, vp, vp, u64, ctypes.POINTER(u64)]),
        "igc_graph_count": (i, [vp, u8p, u32, u32, u32, ctypes.POINTER(u64)]),
        "igc_graph_build_batch": (i, [vp, ctypes.POINTER(_Set), u32, ctypes.POINTER(_Csr)]),
        "igc_kmeans": (i, [vp, u8p, u32, u32, u32, u32, u64, u32, vp, vp]),
        "igc_kmeans_seeds": (i, [u32, u32, u64, vp]),
        "igc_refine": (i, [vp, u8p, u32, u32, u32, u32, u32, u32, u32, u64, u32, vp, ctypes.POINTER(u32),
                           ctypes.POINTER(_RefineStats)]),
        "igc_graph_build_batch_device": (i, [vp, ctypes.POINTER(_Set), u32, ctypes.POINTER(_Csr), vp, vp]),
        "igc_csr_to_host": (i, [vp, ctypes.POINTER(_Csr), u32, ctypes.POINTER(_Csr)]),
        "igc_csr_free": (i, [ctypes.POINTER(_Csr)]),
        "igc_csr_reorder": (i, [ctypes.POINTER(_Csr), u32]),
        "igc_greedy": (i, [ctypes.POINTER(_Csr), vp, ctypes.POINTER(u32)]),
        "igc_greedy_mt": (i, [ctypes.POINTER(_Csr), vp, ctypes.POINTER(u32), u32]),
        "igc_write_pair_file": (i, [ctypes.POINTER(_Csr), ctypes.c_char_p]),
        "igc_write_codebook": (i, [ctypes.c_char_p, u8p, u32, vp, u32, ctypes.POINTER(_CodebookInfo),
                                   ctypes.c_char_p, ctypes.c_size_t]),
        "igc_server_create": (i, [i, u64, u32, u32, ctypes.POINTER(vp)]),
        "igc_server_submit": (i, [vp, u8p, u32, u32, u32, ctypes.POINTER(u64)]),
        "igc_server_submit_refine": (i, [vp, u8p, u32, u32, u32, u32, u32, u32, u32, u64, ctypes.POINTER(u64)]),
        "igc_server_wait": (i, [vp, u64, ctypes.c_int64, ctypes.POINTER(_JobResult)]),
        "igc_server_poll": (i, [vp, u64, ctypes.POINTER(_JobResult)]),
        "igc_job_free": (i, [ctypes.POINTER(_JobResult)]),
        "igc_server_clock": (ctypes.c_double, [vp]),
        "igc_server_stats": (i, [vp, ctypes.POINTER(_ServerInfo)]),
        "igc_server_destroy": (i, [vp]),
        "igc_autotune": (i, [vp, u32, u32, u32, u32, ctypes.POINTER(_TuneEntry), u32, ctypes.POINTER(u32)]),
        "igc_tuned_shape": (i, [vp, u32, u32, u32, ctypes.POINTER(i), ctypes.POINTER(u32)]),
        "igc_pair_plan_of": (i, [vp, u32, u32, u32, ctypes.POINTER(_PairPlan)]),
        "igc_ctx_set_prefilter": (i, [vp, i]),
        "igc_ctx_set_predicate": (i, [vp, i]),
        "igc_ctx_set_csr_form": (i, [vp, i]),
        "igc_ctx_set_edge_cap": (i, [vp, u64]),
        "igc_ctx_set_graph_replay": (i, [vp, i]),
        "igc_assign_fp32": (i, [vp, vp, u32, u32, vp, u32, vp]),
        "igc_gen_linear": (i, [vp, vp, u32, u32, vp, ctypes.POINTER(_Filter), ctypes.POINTER(_Cands)]),
        "igc_gen_vt": (i, [vp, u32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_Filter), ctypes.POINTER(_Cands)]),
        "igc_gen_diff_vt": (i, [vp, u32, ctypes.c_int32, ctypes.POINTER(_Filter), ctypes.POINTER(_Cands)]),
        "igc_gen_all": (i, [vp, u32, ctypes.POINTER(_Filter), ctypes.POINTER(_Cands)]),
        "igc_candidates_to_host": (i, [vp, ctypes.POINTER(_Cands), vp, u64]),
        "igc_last_error": (ctypes.c_char_p, []),
        "igc_backend": (ctypes.c_char_p, []),
        "igc_version": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != IGC_OK:
        raise IGCError(rc, lib().igc_last_error().decode(errors="replace"))


def backend() -> str:
    return lib().igc_backend().decode()


# ---------------------------------------------------------------------------
# candidate representation (VecToStr, src/Utils.cpp:601-609: symbols 0..3 as '0'..'3')

_ACGT = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}


def to_symbols(candidates) -> np.ndarray:
    """N x n uint8 array of symbol codes 0..3 from strings of '0'..'3' or 'ACGT',
    or from an existing integer array."""
    if isinstance(candidates, np.ndarray):
        a = np.ascontiguousarray(candidates, dtype=np.uint8)
        if a.ndim != 2:
            raise ValueError("symbol array must be 2-D (N, n)")
        return a
    cands = list(candidates)
    if not cands:
        return np.zeros((0, 1), dtype=np.uint8)
    n = len(cands[0])
    if any(len(c) != n for c in cands):
        raise ValueError("candidates must all have the same length")
    raw = np.frombuffer("".join(cands).encode("ascii"), dtype=np.uint8).reshape(len(cands), n)
    out = raw - ord("0")
    if (out > 3).any():
        lut = np.full(256, 255, dtype=np.uint8)
        for k, v in _ACGT.items():
            lut[k] = v
        for k in range(4):
            lut[ord("0") + k] = k
        out = lut[raw]
    if (out > 3).any():
        raise ValueError("candidates must use symbols 0-3 or A,C,G,T")
    return np.ascontiguousarray(out, dtype=np.uint8)


def to_strings(syms: np.ndarray) -> list:
    s = np.ascontiguousarray(syms, dtype=np.uint8) + ord("0")
    n = s.shape[1]
    b = s.tobytes().decode("ascii")
    return [b[k * n:(k + 1) * n] for k in range(s.shape[0])]


# ---------------------------------------------------------------------------
# graph objects


@dataclass
class AdjList:
    """Symmetric conflict graph in CSR form (the reference's AdjList,
    include/SparseMat.hpp:42-194, as flat arrays)."""

    row_ptr: np.ndarray  # uint64 [N+1]
    col: np.ndarray      # uint32 [2|E|], or [|E|] for an upper-triangle CSR
    upper: bool = False  # rows list only neighbours above the row (IGC_UPPER_TRIANGLE)

    @property
    def n_nodes(self) -> int:
        return int(self.row_ptr.size - 1)

    @property
    def n_edges(self) -> int:
        return int(self.col.size if self.upper else self.col.size // 2)

    @property
    def matrix_ones(self) -> int:
        """The reference's matrixOnesNum / "Number of Ones in Matrix" = 2|E|."""
        return 2 * self.n_edges

    def neighbors(self, i: int) -> np.ndarray:
        return self.col[int(self.row_ptr[i]):int(self.row_ptr[i + 1])]

    def _as_c(self) -> _Csr:
        c = _Csr()
        c.n_nodes = self.n_nodes
        c.location = IGC_ON_HOST
        c.n_edges = self.n_edges
        self._rp = np.ascontiguousarray(self.row_ptr, dtype=np.uint64)
        self._col = np.ascontiguousarray(self.col, dtype=np.uint32)
        if self._col.size == 0:
            self._col = np.zeros(1, dtype=np.uint32)
        c.row_ptr = self._rp.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        c.col = self._col.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        return c


class _CsrOwner:
    """Holds a library-allocated host CSR; releases it (igc_csr_free) when the
    last numpy view of its arrays is gone."""

    def __init__(self, c: _Csr):
        self.c = c

    def __del__(self):
        try:
            if _lib is not None:
                _lib.igc_csr_free(ctypes.byref(self.c))
        except Exception:  # interpreter shutdown: module globals may already be gone
            pass


def _view(ptr, ctype, count: int, owner: _CsrOwner, dtype) -> np.ndarray:
    addr = ctypes.cast(ptr, ctypes.c_void_p).value
    if not addr:
        return np.zeros(count, dtype=dtype)
    buf = (ctype * max(count, 1)).from_address(addr)
    buf._owner = owner  # the arrays keep the CSR alive; no copy of the library's arrays
    return np.frombuffer(buf, dtype=dtype)[:count]


def _adj_from_c(c: _Csr) -> AdjList:
    N = int(c.n_nodes)
    owner = _CsrOwner(c)
    rp = _view(c.row_ptr, ctypes.c_uint64, N + 1, owner, np.uint64)
    ents = int(rp[N])
    col = _view(c.col, ctypes.c_uint32, ents, owner, np.uint32)
    upper = ents != 2 * int(c.n_edges)
    return AdjList(rp, col, upper)


@dataclass
class Timing:
    pack_ms: float
    pairs_ms: float
    csr_ms: float
    total_ms: float
    pairs: int
    n_edges: int
    launches: int
    retries: int


@dataclass
class DeviceCandidates:
    """Candidates a generator left in device memory (owned by the context;
    valid until its next generator call), in the reference generator's order."""

    d_syms: int  # device address of count x n symbols
    count: int
    enumerated: int
    n: int
    ms: float
    _raw: _Cands = None


@dataclass
class DeviceCSR:
    """CSR left in device memory by Context.graph_build_device (owned by the
    context; valid until its next build)."""

    n_nodes: int
    n_edges: int
    row_ptr: int  # device address
    col: int      # device address
    _raw: _Csr = None


class Context:
    """One GPU, one HIP stream, cached device buffers (an igc_ctx)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().igc_ctx_create(int(device), ctypes.byref(h)))
        self._h = h
        self._fin = weakref.finalize(self, lib().igc_ctx_destroy, h)

    def close(self) -> None:
        self._fin()

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def set_stream(self, stream_ptr: int | None) -> None:
        _check(lib().igc_ctx_set_stream(self._h, ctypes.c_void_p(stream_ptr or 0)))

    @property
    def stream(self) -> int:
        return int(lib().igc_ctx_stream(self._h) or 0)

    def graph_build(self, syms, min_ed: int, flags: int = 0) -> AdjList:
        s = to_symbols(syms)
        N, n = s.shape
        c = _Csr()
        _check(lib().igc_graph_build(self._h, s.ctypes.data, N, n, int(min_ed), int(flags), ctypes.byref(c)))
        return _adj_from_c(c)

    def graph_build_to_device(self, syms, min_ed: int, n: int | None = None, N: int | None = None,
                              flags: int = 0) -> DeviceCSR:
        """Symbols from host memory (a numpy array, or the address of N x n
        bytes, e.g. a pinned torch tensor's data_ptr()), CSR left in HBM
        (igc_graph_build with IGC_RESULT_ON_DEVICE)."""
        if isinstance(syms, int):
            ptr, N, n = syms, int(N), int(n)
        else:
            s = to_symbols(syms)
            N, n = s.shape
            ptr = s.ctypes.data
        c = _Csr()
        _check(lib().igc_graph_build(self._h, ctypes.c_void_p(ptr), N, n, int(min_ed),
                                     int(flags) | IGC_RESULT_ON_DEVICE, ctypes.byref(c)))
        return DeviceCSR(int(c.n_nodes), int(c.n_edges), ctypes.cast(c.row_ptr, ctypes.c_void_p).value or 0,
                         ctypes.cast(c.col, ctypes.c_void_p).value or 0, c)

    def graph_build_device(self, d_syms_ptr: int, N: int, n: int, min_ed: int) -> DeviceCSR:
        c = _Csr()
        _check(lib().igc_graph_build_device(self._h, ctypes.c_void_p(d_syms_ptr), int(N), int(n), int(min_ed),
                                            ctypes.byref(c)))
        return DeviceCSR(int(c.n_nodes), int(c.n_edges), ctypes.cast(c.row_ptr, ctypes.c_void_p).value or 0,
                         ctypes.cast(c.col, ctypes.c_void_p).value or 0, c)

    def graph_build_device_noalloc(self, d_syms_ptr: int, N: int, n: int, min_ed: int, d_row_ptr: int,
                                   d_col: int, col_cap: int) -> int:
        """igc_graph_build_device into caller device buffers (row_ptr N+1 u64,
        col col_cap u32) with no allocation; returns the edge count."""
        e = ctypes.c_uint64()
        _check(lib().igc_graph_build_device_noalloc(self._h, ctypes.c_void_p(d_syms_ptr), int(N), int(n),
                                                    int(min_ed), ctypes.c_void_p(d_row_ptr), ctypes.c_void_p(d_col),
                                                    int(col_cap), ctypes.byref(e)))
        return int(e.value)

    # -- candidate generators (the reference's CandidateGenerator methods + applyFilters)

    @staticmethod
    def _filter(max_run, min_gc, max_gc):
        return _Filter(int(max_run), 0, float(min_gc), float(max_gc))

    def _cands(self, c: _Cands) -> DeviceCandidates:
        return DeviceCandidates(int(c.d_syms or 0), int(c.count), int(c.enumerated), int(c.n), float(c.ms), c)

    def gen_linear(self, G, bias=None, max_run: int = 0, min_gc: float = 0.0, max_gc: float = 0.0) -> DeviceCandidates:
        """LinearCode candidates for generator matrix G (k x n over GF(4), after
        permutations) and bias, filtered, left on the device."""
        G = np.ascontiguousarray(G, dtype=np.uint8)
        if G.ndim != 2:
            raise ValueError("generator matrix must be 2-D (k, n)")
        k, n = G.shape
        b = None if bias is None else np.ascontiguousarray(bias, dtype=np.uint8)
        if b is not None and b.shape != (n,):
            raise ValueError(f"bias must have {n} entries")
        c = _Cands()
        f = self._filter(max_run, min_gc, max_gc)
        _check(lib().igc_gen_linear(self._h, G.ctypes.data, k, n, None if b is None else b.ctypes.data,
                                    ctypes.byref(f), ctypes.byref(c)))
        return self._cands(c)

    def gen_vt(self, n: int, a: int, b: int, max_run: int = 0, min_gc: float = 0.0, max_gc: float = 0.0):
        c = _Cands()
        f = self._filter(max_run, min_gc, max_gc)
        _check(lib().igc_gen_vt(self._h, int(n), int(a), int(b), ctypes.byref(f), ctypes.byref(c)))
        return self._cands(c)

    def gen_diff_vt(self, n: int, syndrome: int, max_run: int = 0, min_gc: float = 0.0, max_gc: float = 0.0):
        c = _Cands()
        f = self._filter(max_run, min_gc, max_gc)
        _check(lib().igc_gen_diff_vt(self._h, int(n), int(syndrome), ctypes.byref(f), ctypes.byref(c)))
        return self._cands(c)

    def gen_all(self, n: int, max_run: int = 0, min_gc: float = 0.0, max_gc: float = 0.0):
        c = _Cands()
        f = self._filter(max_run, min_gc, max_gc)
        _check(lib().igc_gen_all(self._h, int(n), ctypes.byref(f), ctypes.byref(c)))
        return self._cands(c)

    def candidates_to_host(self, c: DeviceCandidates) -> np.ndarray:
        out = np.zeros((max(c.count, 1), c.n), dtype=np.uint8)
        _check(lib().igc_candidates_to_host(self._h, ctypes.byref(c._raw), out.ctypes.data, c.count))
        return out[: c.count]

    def graph_build_candidates(self, c: DeviceCandidates, min_ed: int, flags: int = 0) -> AdjList:
        """Conflict graph of generated candidates without a host round trip of
        the strings: the build reads them where the generator left them."""
        d = self.graph_build_device(c.d_syms, c.count, c.n, min_ed)
        return self.device_csr_to_host(d, flags)

    def device_csr_to_host(self, dcsr: DeviceCSR, flags: int = 0) -> AdjList:
        c = _Csr()
        _check(lib().igc_csr_to_host(self._h, ctypes.byref(dcsr._raw), int(flags), ctypes.byref(c)))
        return _adj_from_c(c)

    def graph_build_batch(self, sets, flags: int = 0) -> list:
        """Conflict graphs of many independent sets [(candidates, min_ed), ...] in
        one persistent launch; one AdjList per set, in order."""
        arrs = [to_symbols(c) for c, _ in sets]
        cs = (_Set * max(len(sets), 1))()
        for k, (a, (_, ed)) in enumerate(zip(arrs, sets)):
            cs[k] = _Set(a.ctypes.data if a.size else None, a.shape[0], a.shape[1], int(ed), int(flags))
        outs = (_Csr * max(len(sets), 1))()
        _check(lib().igc_graph_build_batch(self._h, cs, len(sets), outs))
        return [_adj_from_c(outs[k]) for k in range(len(sets))]

    def graph_build_batch_device(self, dev_sets):
        """dev_sets: [(device_ptr, N, n, min_ed), ...].  Returns (union DeviceCSR,
        node_off list, per-set edge counts); the CSR stays in HBM."""
        m = len(dev_sets)
        cs = (_Set * m)(*[_Set(p, N, n, ed, 0) for p, N, n, ed in dev_sets])
        off = np.zeros(m + 1, dtype=np.uint32)
        edges = np.zeros(m, dtype=np.uint64)
        c = _Csr()
        _check(lib().igc_graph_build_batch_device(self._h, cs, m, ctypes.byref(c), off.ctypes.data,
                                                  edges.ctypes.data))
        d = DeviceCSR(int(c.n_nodes), int(c.n_edges), ctypes.cast(c.row_ptr, ctypes.c_void_p).value or 0,
                      ctypes.cast(c.col, ctypes.c_void_p).value or 0, c)
        return d, off, edges

    def kmeans(self, syms, k: int, iters: int, seed: int, eager: bool = False):
        """FP32 K-means on one-hot features: (assign int32[N], centroids float32[k, 4n])."""
        s = to_symbols(syms)
        N, n = s.shape
        assign = np.zeros(max(N, 1), dtype=np.int32)
        C = np.zeros((k, 4 * n), dtype=np.float32)
        _check(lib().igc_kmeans(self._h, s.ctypes.data, N, n, int(k), int(iters), int(seed),
                                IGC_KMEANS_EAGER if eager else 0, assign.ctypes.data, C.ctypes.data))
        return assign[:N], C

    def refine(self, syms, min_ed: int, k: int, kmeans_iters: int = 10, max_rounds: int = 5, converge_n: int = 3,
               seed: int = 1, host_threads: int = 8):
        """Clustered refinement loop; returns (final codebook indices, stats dict)."""
        s = to_symbols(syms)
        N, n = s.shape
        out = np.zeros(max(N, 1), dtype=np.uint32)
        cnt = ctypes.c_uint32()
        st = _RefineStats()
        _check(lib().igc_refine(self._h, s.ctypes.data, N, n, int(min_ed), int(k), int(kmeans_iters),
                                int(max_rounds), int(converge_n), int(seed), int(host_threads), out.ctypes.data,
                                ctypes.byref(cnt), ctypes.byref(st)))
        rounds = []
        for r in range(min(st.rounds, IGC_REFINE_MAX_ROUNDS)):
            x = st.round[r]
            rounds.append({"candidates": x.candidates, "clusters": x.clusters, "codewords": x.codewords,
                           "pairs": int(x.pairs), "kmeans_ms": x.kmeans_ms, "pairs_ms": x.pairs_ms,
                           "cluster_s": x.cluster_s, "graph_s": x.graph_s, "greedy_s": x.greedy_s})
        return out[: cnt.value].copy(), {"rounds": rounds, "converged": bool(st.converged)}

    def autotune(self, n: int, min_ed: int, N: int, reps: int = 3) -> list:
        """Times every tile shape for (n, min_ed, N) and adopts the fastest; returns
        [{cpl, q, ms, gpairs_s}, ...]."""
        ents = (_TuneEntry * 32)()
        cnt = ctypes.c_uint32()
        _check(lib().igc_autotune(self._h, int(n), int(min_ed), int(N), int(reps), ents, 32, ctypes.byref(cnt)))
        return [{"cpl": e.cpl, "q": e.q, "ms": e.ms, "gpairs_s": e.gpairs_s} for e in ents[: cnt.value]]

    def tuned_shape(self, n: int, min_ed: int, N: int):
        cpl, q = ctypes.c_int(), ctypes.c_uint32()
        _check(lib().igc_tuned_shape(self._h, int(n), int(min_ed), int(N), ctypes.byref(cpl), ctypes.byref(q)))
        return cpl.value, q.value

    def assign_fp32(self, syms, centroids) -> np.ndarray:
        """Nearest one-hot FP32 centroid of every string (int32[N]); centroids float32[k, 4n]."""
        s = to_symbols(syms)
        C = np.ascontiguousarray(centroids, dtype=np.float32)
        N, n = s.shape
        if C.ndim != 2 or C.shape[1] != 4 * n:
            raise ValueError(f"centroids must be [k, {4 * n}]")
        out = np.zeros(N, dtype=np.int32)
        _check(lib().igc_assign_fp32(self._h, s.ctypes.data, N, n, C.ctypes.data, C.shape[0], out.ctypes.data))
        return out

    def set_prefilter(self, mode) -> None:
        """Exact row skipping by half-string symbol counts (same graph either way):
        False/IGC_PREFILTER_OFF, True/IGC_PREFILTER_AUTO (default: sets of at least
        IGC_PREFILTER_MIN_N strings) or IGC_PREFILTER_ALWAYS."""
        m = int(mode) if not isinstance(mode, bool) else (IGC_PREFILTER_AUTO if mode else IGC_PREFILTER_OFF)
        _check(lib().igc_ctx_set_prefilter(self._h, m))

    def set_predicate(self, predicate: int) -> None:
        """Pair predicate of the kernels: IGC_PRED_AUTO (default, the cheaper for
        (n, editDist)), IGC_PRED_LANDAU_VISHKIN (where editDist - 1 <= 13) or
        IGC_PRED_MYERS; the same graph either way."""
        _check(lib().igc_ctx_set_predicate(self._h, int(predicate)))

    def set_csr_form(self, form: int) -> None:
        """CSR assembly: IGC_CSR_AUTO (split pass + 32-bit radix passes where they
        fit) or IGC_CSR_WIDE_KEYS (64-bit keys); the same CSR either way."""
        _check(lib().igc_ctx_set_csr_form(self._h, int(form)))

    def set_graph_replay(self, on: bool) -> None:
        """refine replays its captured round graph (True, default) or launches
        the same kernels directly (False); the same results."""
        _check(lib().igc_ctx_set_graph_replay(self._h, 1 if on else 0))

    def set_edge_cap(self, edges: int) -> None:
        """First edge buffer of a build, in edges (0: the density estimate); a
        graph with more edges regrows it to the exact count and replays the pair
        kernels (last_timing().retries), with the same result."""
        _check(lib().igc_ctx_set_edge_cap(self._h, int(edges)))

    def pair_plan(self, n: int, min_ed: int, N: int) -> dict:
        """The pair launch this context makes for N strings at (n, min_ed)."""
        p = _PairPlan()
        _check(lib().igc_pair_plan_of(self._h, int(n), int(min_ed), int(N), ctypes.byref(p)))
        return {"predicate": PREDICATES[p.predicate], "k": p.k, "lane_cols": p.lane_cols, "slots": p.slots,
                "q": p.q, "tile_cols": p.tile_cols, "tile_rows": p.tile_rows, "tiles": int(p.tiles),
                "row_filter": bool(p.row_filter)}

    def graph_count(self, syms, min_ed: int) -> int:
        s = to_symbols(syms)
        N, n = s.shape
        out = ctypes.c_uint64()
        _check(lib().igc_graph_count(self._h, s.ctypes.data, N, n, int(min_ed), ctypes.byref(out)))
        return int(out.value)

    def last_timing(self) -> Timing:
        t = _Timing()
        _check(lib().igc_last_timing(self._h, ctypes.byref(t)))
        return Timing(t.pack_ms, t.pairs_ms, t.csr_ms, t.total_ms, int(t.pairs), int(t.n_edges), int(t.launches),
                      int(t.retries))


class Server:
    """Continuous-batching codebook server (igc_server): submit jobs from any
    thread, collect codebooks with their timelines."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, pipelines: int = 2, host_threads: int = 8):
        h = ctypes.c_void_p()
        _check(lib().igc_server_create(int(device), int(hbm_budget_bytes), int(pipelines), int(host_threads),
                                       ctypes.byref(h)))
        self._h = h
        self._fin = weakref.finalize(self, lib().igc_server_destroy, h)

    def close(self) -> None:
        self._fin()

    def clock(self) -> float:
        return float(lib().igc_server_clock(self._h))

    def submit(self, candidates, min_ed: int) -> int:
        s = to_symbols(candidates)
        jid = ctypes.c_uint64()
        _check(lib().igc_server_submit(self._h, s.ctypes.data if s.size else None, s.shape[0], s.shape[1],
                                       int(min_ed), ctypes.byref(jid)))
        return int(jid.value)

    def submit_refine(self, candidates, min_ed: int, k: int, kmeans_iters: int = 10, max_rounds: int = 5,
                      converge_n: int = 3, seed: int = 1) -> int:
        """A refinement job (the clustered loop of Context.refine); its result's
        order is the final codebook."""
        s = to_symbols(candidates)
        jid = ctypes.c_uint64()
        _check(lib().igc_server_submit_refine(self._h, s.ctypes.data if s.size else None, s.shape[0], s.shape[1],
                                              int(min_ed), int(k), int(kmeans_iters), int(max_rounds),
                                              int(converge_n), int(seed), ctypes.byref(jid)))
        return int(jid.value)

    def wait(self, job_id: int, timeout_ms: int = -1):
        """dict with status, order (np.uint32), n_edges, timeline; None on timeout."""
        r = _JobResult()
        rc = lib().igc_server_wait(self._h, int(job_id), int(timeout_ms), ctypes.byref(r))
        if rc == IGC_ERR_BUSY:
            return None
        _check(rc)
        order = (np.ctypeslib.as_array(r.order, shape=(r.codewords,)).copy() if r.codewords
                 else np.zeros(0, dtype=np.uint32))
        lib().igc_job_free(ctypes.byref(r))
        err = lib().igc_last_error().decode() if r.status else ""
        return {"id": int(r.id), "status": int(r.status), "error": err, "order": order, "n_nodes": int(r.n_nodes),
                "n_edges": int(r.n_edges), "submit_s": r.submit_s, "admit_s": r.admit_s,
                "gpu_done_s": r.gpu_done_s, "greedy_start_s": r.greedy_start_s, "done_s": r.done_s,
                "batch": int(r.batch), "batch_jobs": int(r.batch_jobs)}

    def stats(self) -> dict:
        x = _ServerInfo()
        _check(lib().igc_server_stats(self._h, ctypes.byref(x)))
        return {k: getattr(x, k) for k, _ in _ServerInfo._fields_}


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ---------------------------------------------------------------------------
# reference-shaped host API


def fill_adj_list_gpu(candidates, min_ed: int, ctx: Context | None = None, threads_order: int = 1):
    """FillAdjListGPU (src/SparseMat.cpp:454-534): the conflict graph of the
    candidates (edge iff edit distance < min_ed).  Returns (AdjList, matrix_ones)
    with matrix_ones = 2|E| as the reference reports it.  threads_order=T
    reproduces the neighbour order of the reference's T-thread CPU builder."""
    ctx = ctx or default_context()
    flags = order_threads(threads_order) if threads_order > 1 else 0
    adj = ctx.graph_build(candidates, min_ed, flags)
    return adj, adj.matrix_ones


def reorder(adj: AdjList, flags: int) -> AdjList:
    """Copy of an ascending AdjList with neighbour lists in the order `flags` names."""
    out = AdjList(adj.row_ptr.copy(), adj.col.copy())
    c = out._as_c()
    _check(lib().igc_csr_reorder(ctypes.byref(c), int(flags)))
    out.col = out._col[: out.col.size].copy()
    return out


def greedy_order(adj: AdjList, threads: int = 0) -> np.ndarray:
    """Node ids of the min-degree greedy independent set, in selection order
    (AdjList::FindMinDel loop of Codebook, src/SparseMat.cpp:698-729).
    threads: 0 = automatic (one below 2^24 edges, else three: the moving
    thread and two that scan ahead of it); any count gives the same
    selection."""
    c = adj._as_c()
    order = np.zeros(max(adj.n_nodes, 1), dtype=np.uint32)
    cnt = ctypes.c_uint32()
    _check(lib().igc_greedy_mt(ctypes.byref(c), order.ctypes.data, ctypes.byref(cnt), int(threads)))
    return order[: cnt.value].copy()


def edges_sha256(adj: AdjList) -> str:
    """SHA-256 of the graph's edge set as LE uint32 pairs (i, j), i < j, sorted
    by (i, j) -- the reference's pair list in a canonical order.  Needs
    ascending neighbour lists (the default order, or an upper-triangle CSR)."""
    import hashlib
    rp = adj.row_ptr.astype(np.int64)
    N = rp.size - 1
    rows = np.repeat(np.arange(N, dtype=np.uint32), np.diff(rp))
    keep = adj.col > rows
    pairs = np.empty((int(keep.sum()), 2), dtype="<u4")
    pairs[:, 0] = rows[keep]
    pairs[:, 1] = adj.col[keep]
    return hashlib.sha256(pairs.tobytes()).hexdigest()


def codebook_sha256(candidates, order) -> str:
    """SHA-256 of the codebook as its words ('0'..'3') joined by newlines, in
    selection order."""
    import hashlib
    syms = to_symbols(candidates)
    words = np.ascontiguousarray(syms[np.asarray(order, dtype=np.int64)], dtype=np.uint8) + ord("0")
    n = syms.shape[1]
    lines = np.full((words.shape[0], n + 1), ord("\n"), dtype=np.uint8)
    lines[:, :n] = words
    return hashlib.sha256(lines.tobytes()[:-1] if words.shape[0] else b"").hexdigest()


def codebook(adj: AdjList, candidates) -> list:
    """Codebook (src/SparseMat.cpp:679-731): codewords in selection order."""
    syms = to_symbols(candidates)
    return to_strings(syms[greedy_order(adj)])


def solve_independent_set(candidates, min_ed: int, ctx: Context | None = None) -> list:
    """SolveIndependentSet (src/SparseMat.cpp:784-851)."""
    syms = to_symbols(candidates)
    if syms.shape[0] == 0:
        return []
    if syms.shape[0] == 1:
        return to_strings(syms)
    adj, _ = fill_adj_list_gpu(syms, min_ed, ctx)
    return to_strings(syms[greedy_order(adj)])


def verify_dist(codewords, min_ed: int, ctx: Context | None = None) -> bool:
    """VerifyDist (src/Utils.cpp:203-309): True iff no pair is closer than min_ed."""
    syms = to_symbols(codewords)
    if syms.shape[0] < 2:
        return True
    ctx = ctx or default_context()
    return ctx.graph_count(syms, min_ed) == 0


def write_pair_file(adj: AdjList, path: str) -> None:
    """The reference's LE int32 (u, v) u<v edge file (src/gpu_graph_generator.py:57-77)."""
    c = adj._as_c()
    _check(lib().igc_write_pair_file(ctypes.byref(c), path.encode()))


def write_codebook(directory: str, candidates, order: Sequence[int], *, code_len: int, min_ed: int,
                   max_run: int = 0, min_gc: float = 0.0, max_gc: float = 0.0, method_line: str | None = None,
                   candidate_num: int = 0, matrix_ones: int = 0, threads: int = 1, candidates_s: float = 0.0,
                   fill_s: float = 0.0, process_s: float = 0.0, overall_s: float = 0.0, cluster_k: int = -1,
                   cluster_iterations: int = -1, convergence_iterations: int = 3,
                   cluster_method: str | None = None) -> str:
    """ToFile (src/Utils.cpp:381-418): writes CodeSize-..._CodeLen-.._MinED-..txt and
    returns its path.  cluster_k > 0 writes the clustered run's header
    (fill_s = total clustering time, process_s = total solving time)."""
    syms = to_symbols(candidates)
    order = np.ascontiguousarray(order, dtype=np.uint32)
    info = _CodebookInfo(code_len, min_ed, max_run, min_gc, max_gc,
                         method_line.encode() if method_line else None, candidate_num, matrix_ones, threads,
                         candidates_s, fill_s, process_s, overall_s, cluster_k, cluster_iterations,
                         convergence_iterations, cluster_method.encode() if cluster_method else None)
    name = ctypes.create_string_buffer(128)
    _check(lib().igc_write_codebook(directory.encode(), syms.ctypes.data, syms.shape[1], order.ctypes.data,
                                    order.size, ctypes.byref(info), name, 128))
    return os.path.join(directory, name.value.decode())


def kmeans_seeds(N: int, k: int, seed: int) -> np.ndarray:
    out = np.zeros(k, dtype=np.uint32)
    _check(lib().igc_kmeans_seeds(int(N), int(k), int(seed), out.ctypes.data))
    return out
