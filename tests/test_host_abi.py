"""CPU-side checks of libigc.so: it loads, exports every symbol include/igc.h
declares, refuses to run without a gfx950 device (no CPU fallback), and its
host-only parts (greedy, neighbour order, writers) match the reference's
golden outputs.  No GPU compute is called here."""
import os
import re
import struct

import numpy as np
import pytest

import indexgen_amd as igc
from indexgen_amd import workloads as W
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"IGC_API\s+[\w\s\*]*?\b(igc_\w+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = igc.lib()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_no_cpu_fallback_without_gpu():
    if igc.backend() != "none":
        pytest.skip("a GPU is visible")
    with pytest.raises(igc.IGCError) as e:
        igc.Context(0)
    assert e.value.code == igc.IGC_ERR_DEVICE


def test_host_greedy_matches_reference(graphs_golden):
    for s in graphs_golden:
        for T, (rp, col, words) in s["T"].items():
            adj = igc.AdjList(rp, col)
            assert np.array_equal(s["syms"][igc.greedy_order(adj)], words)


def test_thread_order_reorder_matches_reference(graphs_golden):
    for s in graphs_golden:
        rp1, col1, _ = s["T"][1]
        rp4, col4, _ = s["T"][4]
        out = igc.reorder(igc.AdjList(rp1, col1), igc.order_threads(4))
        assert np.array_equal(out.row_ptr, rp4)
        assert np.array_equal(out.col, col4)


def test_greedy_matches_oracle_on_larger_graphs():
    rng = np.random.default_rng(11)
    for n, ed in [(8, 3), (10, 4), (12, 5)]:
        syms = W.dedup_first(rng.integers(0, 4, size=(3000, n), dtype=np.uint8))
        for T in (1, 3):
            rp, col = O.graph_csr(syms, ed, T=T)
            assert np.array_equal(igc.greedy_order(igc.AdjList(rp, col)).astype(np.int32), O.greedy(rp, col))


def _random_graph(rng, N: int, p: float, cliques: int = 0) -> "igc.AdjList":
    """Symmetric CSR with ascending rows: G(N, p), plus `cliques` disjoint
    cliques of 20 nodes (many equal degrees, so ties decide picks)."""
    iu, ju = np.triu_indices(N, 1)
    keep = rng.random(iu.size) < p
    i, j = iu[keep], ju[keep]
    if cliques:
        members = rng.permutation(N)[: 20 * cliques].reshape(cliques, 20)
        ci, cj = np.triu_indices(20, 1)
        i = np.concatenate([i, np.minimum(members[:, ci], members[:, cj]).ravel()])
        j = np.concatenate([j, np.maximum(members[:, ci], members[:, cj]).ravel()])
        key = np.unique(i.astype(np.int64) * N + j)
        i, j = key // N, key % N
    src = np.concatenate([i, j])
    dst = np.concatenate([j, i])
    o = np.lexsort((dst, src))
    rp = np.zeros(N + 1, np.uint64)
    np.add.at(rp, src[o] + 1, 1)
    return igc.AdjList(np.cumsum(rp).astype(np.uint64), dst[o].astype(np.uint32))


def test_greedy_scanning_ahead_equals_one_thread():
    """The greedy with helper threads (for picks that scan >= 2^15 neighbours,
    the helpers find each ball member's active neighbours ahead of the moves)
    selects exactly what one thread selects, also with rows in the reference's
    T-thread order instead of ascending, and on a conflict graph against the
    oracle's greedy."""
    rng = np.random.default_rng(21)
    graphs = [_random_graph(rng, 6000, 0.15), _random_graph(rng, 5000, 0.05, cliques=120),
              _random_graph(rng, 3000, 0.6)]
    graphs.append(igc.reorder(graphs[1], igc.order_threads(3)))
    for adj in graphs:
        one = igc.greedy_order(adj, threads=1)
        for T in (2, 3, 8, 16):
            assert np.array_equal(igc.greedy_order(adj, threads=T), one), T
    syms = W.dedup_first(rng.integers(0, 4, size=(6000, 9), dtype=np.uint8))
    rp, col = O.graph_csr(syms, 5)
    assert rp[-1] > 2_000_000  # dense: most picks are split
    want = O.greedy(rp, col)
    for T in (1, 4, 7):
        assert np.array_equal(igc.greedy_order(igc.AdjList(rp, col), threads=T).astype(np.int32), want)


def test_pair_file_format(tmp_path):
    rp = np.array([0, 2, 3, 4, 5], np.uint64)
    col = np.array([1, 3, 0, 3, 0], np.uint32)[:5]
    # 4 nodes: 0-1, 0-3, 2-3 ... build a consistent symmetric CSR instead
    edges = [(0, 1), (0, 3), (2, 3)]
    adj = [[] for _ in range(4)]
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    rp = np.array([0] + list(np.cumsum([len(a) for a in adj])), np.uint64)
    col = np.array(sum((sorted(a) for a in adj), []), np.uint32)
    p = tmp_path / "e.bin"
    igc.write_pair_file(igc.AdjList(rp, col), str(p))
    data = p.read_bytes()
    assert data == b"".join(struct.pack("<ii", u, v) for u, v in edges)


def test_codebook_file_format(tmp_path):
    syms = np.array([[0, 1, 2, 3], [3, 2, 1, 0], [1, 1, 2, 2]], np.uint8)
    path = igc.write_codebook(str(tmp_path), syms, [2, 0], code_len=4, min_ed=3, max_run=3, min_gc=0.3,
                              max_gc=0.7, candidate_num=3, matrix_ones=2, threads=1, fill_s=1.234)
    assert os.path.basename(path) == "CodeSize-0000002_CodeLen-04_MinED-3.txt"
    lines = open(path).read().split("\n")
    assert lines[0] == "--- Global Parameters ---"
    assert lines[1] == "Code Length:\t\t\t4"
    assert lines[2] == "Min Codebook Edit Distance:\t3"
    assert "Min GC Content:\t\t\t0.3" in lines
    assert "Number of Ones in Matrix:\t2" in lines
    assert "Fill Adjacency List Time:\t1.23\tseconds" in lines
    sep = lines.index("=========================================== ")
    assert lines[sep + 1:sep + 3] == ["1122", "0123"]


def test_symbol_conversion_round_trip():
    s = ["0123", "3210", "ACGT"]
    syms = igc.to_symbols(s)
    assert syms.tolist() == [[0, 1, 2, 3], [3, 2, 1, 0], [0, 1, 2, 3]]
    assert igc.to_strings(syms) == ["0123", "3210", "0123"]
    with pytest.raises(ValueError):
        igc.to_symbols(["01", "012"])
    with pytest.raises(ValueError):
        igc.to_symbols(["0124"])


def test_workload_filters_match_oracle():
    rng = np.random.default_rng(6)
    syms = rng.integers(0, 4, size=(2000, 12), dtype=np.uint8)
    mask = W.filter_mask(syms, 3, 0.3, 0.7)
    assert np.flatnonzero(mask).tolist() == O.filter_indices(syms, 3, 0.3, 0.7).tolist()


def test_workload_shapes():
    c1 = W.plain_small()
    assert c1.shape[1] == 11 and 14000 < c1.shape[0] < 14900
    assert np.unique(c1, axis=0).shape[0] == c1.shape[0]
    g = W.rng(9)
    P, words = W.systematic_gf4_code(g, 13, 8, 3)
    # systematic part is the message; every nonzero codeword has weight >= 3
    assert words.shape == (65536, 13)
    assert (words[1:] != 0).sum(axis=1).min() >= 3
    # linearity: the sum of two codewords is a codeword
    s = set(map(bytes, words))
    a, b = words[123], words[4567]
    assert bytes(W.GF4_ADD[a, b]) in s
    vt = W.filtered_uniform(W.rng(3), 200, 16, predicate=W.vt_syndrome_ok)
    assert W.vt_syndrome_ok(vt).all() and W.filter_mask(vt).all()
