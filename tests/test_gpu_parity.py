"""GPU parity of the conflict-graph build (libigc.so through its C ABI) with the
oracle and the reference's golden outputs.  Integer work: every comparison is
bit-exact (edge sets, CSR arrays, neighbour order, codebooks)."""
import os

import numpy as np
import pytest

import indexgen_amd as igc
from indexgen_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def check_csr_wellformed(adj: igc.AdjList):
    N = adj.n_nodes
    rp = adj.row_ptr.astype(np.int64)
    assert rp[0] == 0 and rp[-1] == adj.col.size and (np.diff(rp) >= 0).all()
    rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(rp))
    col = adj.col.astype(np.int64)
    assert (col < N).all() and (col != rows).all()
    # ascending within rows
    same = rows[1:] == rows[:-1]
    assert (col[1:][same] > col[:-1][same]).all()
    # symmetric: sorted (row, col) equals sorted (col, row)
    a = np.sort(rows * N + col)
    b = np.sort(col * N + rows)
    assert np.array_equal(a, b)


def test_golden_graphs_and_codebooks(gpu_ctx, graphs_golden):
    for k, s in enumerate(graphs_golden):
        banner(f"golden set {k} shape={s['syms'].shape} min_ed={s['min_ed']}")
        for T, (rp, col, words) in s["T"].items():
            flags = igc.order_threads(T) if T > 1 else 0
            adj = gpu_ctx.graph_build(s["syms"], s["min_ed"], flags)
            assert np.array_equal(adj.row_ptr, rp)
            assert np.array_equal(adj.col, col)
            assert np.array_equal(s["syms"][igc.greedy_order(adj)], words)


@pytest.mark.parametrize("n", list(range(1, 33)))
def test_random_sets_every_length_and_threshold(gpu_ctx, n):
    rng = np.random.default_rng(1000 + n)
    N = 700 if n > 3 else 300
    syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
    # plant near copies so every threshold sees edges
    for r in range(0, N // 4):
        src = syms[int(rng.integers(0, N))].copy()
        src[int(rng.integers(0, n))] = int(rng.integers(0, 4))
        syms[r] = src
    for ed in range(0, n + 3):
        banner(f"n={n} min_ed={ed} N={N}")
        adj = gpu_ctx.graph_build(syms, ed)
        rp, col = O.graph_csr(syms, ed)
        assert np.array_equal(adj.row_ptr, rp), (n, ed)
        assert np.array_equal(adj.col, col), (n, ed)


def _near_copies(rng, syms, rows):
    """Rows 0..rows-1 become copies of random rows with 1-3 substitutions or
    indel pairs, so thresholds see edges."""
    N, n = syms.shape
    for r in range(rows):
        src = list(syms[int(rng.integers(0, N))])
        for _ in range(int(rng.integers(1, 4))):
            p = int(rng.integers(0, n))
            if rng.integers(0, 2):
                src[p] = int(rng.integers(0, 4))
            else:
                del src[p]
                src.insert(int(rng.integers(0, n)), int(rng.integers(0, 4)))
        syms[r] = src


@pytest.mark.parametrize("n", [5, 8, 11, 13, 16, 17, 24, 32])
def test_prefilter_sets_match_oracle(gpu_ctx, n):
    """Sets sorted for the row filter (forced below its automatic size; rows
    skipped per column window): same CSR as the oracle and as the unfiltered
    build."""
    rng = np.random.default_rng(3000 + n)
    N = 6000
    syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
    _near_copies(rng, syms, N // 3)
    syms[N // 3: N // 3 + 50] = 0  # one-symbol strings: all counts in one byte
    gpu_ctx.set_prefilter(igc.IGC_PREFILTER_ALWAYS)
    try:
        for ed in sorted({1, 2, 3, 4, 5, n // 2, n - 1, n, n + 1}):
            if ed < 1:
                continue
            banner(f"prefilter n={n} min_ed={ed} N={N}")
            rp, col = O.graph_csr(syms, ed)
            adj = gpu_ctx.graph_build(syms, ed)
            assert np.array_equal(adj.row_ptr, rp), (n, ed)
            assert np.array_equal(adj.col, col), (n, ed)
            assert gpu_ctx.graph_count(syms, ed) == col.size // 2
        gpu_ctx.set_prefilter(False)
        adj_off = gpu_ctx.graph_build(syms, 4)
    finally:
        gpu_ctx.set_prefilter(True)
    rp, col = O.graph_csr(syms, 4)
    assert np.array_equal(adj_off.col, col)


def test_prefilter_modes(gpu_ctx):
    """The automatic mode sorts only sets of at least IGC_PREFILTER_MIN_N
    strings; bad modes are rejected; every mode gives the same graph."""
    assert igc.IGC_PREFILTER_MIN_N == 16384
    with pytest.raises(igc.IGCError):
        gpu_ctx.set_prefilter(3)
    rng = np.random.default_rng(65)
    syms = rng.integers(0, 4, size=(igc.IGC_PREFILTER_MIN_N, 11), dtype=np.uint8)
    _near_copies(rng, syms, 2000)
    graphs = []
    for mode in (igc.IGC_PREFILTER_OFF, igc.IGC_PREFILTER_ALWAYS, igc.IGC_PREFILTER_AUTO):
        gpu_ctx.set_prefilter(mode)
        graphs.append(gpu_ctx.graph_build(syms, 3))
    gpu_ctx.set_prefilter(True)
    for g in graphs[1:]:
        assert np.array_equal(g.row_ptr, graphs[0].row_ptr) and np.array_equal(g.col, graphs[0].col)


def test_edge_cases(gpu_ctx):
    banner("edge cases")
    # empty and single inputs
    for N in (0, 1):
        adj = gpu_ctx.graph_build(np.zeros((N, 5), np.uint8), 3)
        assert adj.n_nodes == N and adj.col.size == 0
    # two strings
    adj = gpu_ctx.graph_build(np.array([[0, 1, 2], [0, 1, 3]], np.uint8), 2)
    assert adj.col.tolist() == [1, 0]
    # all strings equal: complete graph
    syms = np.tile(np.array([[1, 2, 3, 0, 1]], np.uint8), (300, 1))
    adj = gpu_ctx.graph_build(syms, 1)
    assert adj.n_edges == 300 * 299 // 2
    check_csr_wellformed(adj)
    # min_ed 0: no edges; min_ed > n: every pair
    rng = np.random.default_rng(7)
    syms = W.dedup_first(rng.integers(0, 4, size=(500, 6), dtype=np.uint8))
    assert gpu_ctx.graph_build(syms, 0).n_edges == 0
    assert gpu_ctx.graph_build(syms, 7).n_edges == syms.shape[0] * (syms.shape[0] - 1) // 2
    assert gpu_ctx.graph_build(syms, 40).n_edges == syms.shape[0] * (syms.shape[0] - 1) // 2
    # ragged sizes around tile and wave boundaries
    for N in (63, 64, 65, 255, 256, 257, 1023, 1024, 1025, 4097):
        s = rng.integers(0, 4, size=(N, 9), dtype=np.uint8)
        adj = gpu_ctx.graph_build(s, 4)
        rp, col = O.graph_csr(s, 4)
        assert np.array_equal(adj.row_ptr, rp) and np.array_equal(adj.col, col), N
    # the same on 64-bit words, Landau-Vishkin (n=24, K=5) and Myers (n=32, K=20)
    for N, n, ed in ((63, 24, 6), (65, 24, 6), (257, 24, 6), (1025, 24, 6), (64, 32, 21), (1023, 32, 21)):
        s = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        s[: N // 4] = s[N // 4: 2 * (N // 4)]  # duplicates: edges at every threshold
        adj = gpu_ctx.graph_build(s, ed)
        rp, col = O.graph_csr(s, ed)
        assert np.array_equal(adj.row_ptr, rp) and np.array_equal(adj.col, col), (N, n, ed)
    # invalid input is refused, not computed
    with pytest.raises(igc.IGCError) as e:
        gpu_ctx.graph_build(np.array([[0, 4, 1]], np.uint8).repeat(3, 0), 2)
    assert e.value.code == igc.IGC_ERR_INVALID
    with pytest.raises(igc.IGCError):
        gpu_ctx.graph_build(np.zeros((4, 65), np.uint8), 3)
    # strings longer than 32 are binary: a symbol above 1 is refused
    bad = np.zeros((40, 40), np.uint8)
    bad[17, 5] = 2
    with pytest.raises(igc.IGCError) as e:
        gpu_ctx.graph_build(bad, 3)
    assert e.value.code == igc.IGC_ERR_INVALID and "binary" in str(e.value)


def _binary_set(rng, N, n, near):
    """N random binary strings of length n; rows 0..near-1 are copies of random
    rows with 1-6 bit flips or indel pairs."""
    syms = rng.integers(0, 2, size=(N, n), dtype=np.uint8)
    for r in range(near):
        src = list(syms[int(rng.integers(0, N))])
        for _ in range(int(rng.integers(1, 7))):
            p = int(rng.integers(0, n))
            if rng.integers(0, 2):
                src[p] ^= 1
            else:
                del src[p]
                src.insert(int(rng.integers(0, n)), int(rng.integers(0, 2)))
        syms[r] = src
    return syms


@pytest.mark.parametrize("n", [33, 34, 40, 47, 48, 56, 63, 64])
def test_binary_sets_match_oracle(gpu_ctx, n):
    """Binary strings longer than 32 (the reference's isBinary solver, L <= 64):
    Landau-Vishkin on 64-bit words up to K = 13, Myers above, every pair once
    editDist > n; the same CSR as the oracle at every listed threshold."""
    rng = np.random.default_rng(3300 + n)
    N = 1500
    syms = _binary_set(rng, N, n, N // 3)
    syms[N // 2: N // 2 + 6] = 0  # all-zero and all-one strings
    syms[N // 2 + 6: N // 2 + 12] = 1
    for ed in (0, 1, 2, 3, 4, 5, 8, 11, 14, 15, 16, 20, 31, n - 1, n, n + 1):
        banner(f"binary n={n} min_ed={ed} N={N}")
        adj = gpu_ctx.graph_build(syms, ed)
        rp, col = O.graph_csr(syms, ed)
        assert np.array_equal(adj.row_ptr, rp), (n, ed)
        assert np.array_equal(adj.col, col), (n, ed)
        assert gpu_ctx.graph_count(syms, ed) == adj.n_edges


def test_binary_ragged_sizes_and_thread_order(gpu_ctx):
    rng = np.random.default_rng(77)
    for N, n, ed in ((63, 64, 4), (64, 33, 3), (65, 50, 12), (257, 64, 17), (1025, 40, 6), (4097, 64, 9)):
        banner(f"binary N={N} n={n} min_ed={ed}")
        s = _binary_set(rng, N, n, N // 4)
        adj = gpu_ctx.graph_build(s, ed)
        rp, col = O.graph_csr(s, ed)
        assert np.array_equal(adj.row_ptr, rp) and np.array_equal(adj.col, col), (N, n, ed)
        adj4 = gpu_ctx.graph_build(s, ed, igc.order_threads(4))
        rp4, col4 = O.graph_csr(s, ed, T=4)
        assert np.array_equal(adj4.row_ptr, rp4) and np.array_equal(adj4.col, col4), (N, n, ed)
        check_csr_wellformed(adj)


def test_binary_larger_set_properties(gpu_ctx):
    """A 60,000-string binary set at n = 64: the CSR is well formed, the
    sampled rows equal the oracle's neighbour lists, and the greedy codebook
    is independent and maximal under the oracle's predicate."""
    rng = np.random.default_rng(640)
    N, n, ed = 60000, 64, 9
    syms = _binary_set(rng, N, n, N // 10)
    banner(f"binary N={N} n={n} min_ed={ed}")
    adj = gpu_ctx.graph_build(syms, ed)
    check_csr_wellformed(adj)
    for i in rng.integers(0, N, 12):
        i = int(i)
        got = adj.col[int(adj.row_ptr[i]): int(adj.row_ptr[i + 1])]
        assert np.array_equal(got, O.row_neighbors(syms, ed, i)), i
    order = igc.greedy_order(adj)
    rp, col = O.graph_csr(syms[order[:1500]], ed)
    assert rp[-1] == 0  # independent
    assert gpu_ctx.graph_count(syms, ed) == adj.n_edges


def test_binary_sets_in_a_batch(gpu_ctx):
    """Binary sets (n > 32) get their own groups in a batch next to quaternary
    sets; every CSR equals the oracle's and the single-set build's."""
    rng = np.random.default_rng(4040)
    sets = []
    for n, N, ed in [(64, 1200, 5), (40, 900, 3), (33, 700, 14), (48, 500, 20), (64, 300, 70), (36, 2, 2),
                     (50, 1, 3), (40, 0, 3), (13, 800, 4), (24, 600, 6), (64, 1000, 9)]:
        sets.append((_binary_set(rng, N, n, N // 3) if n > 32 else rng.integers(0, 4, size=(N, n), dtype=np.uint8),
                     ed))
    banner(f"batch of {len(sets)} binary and quaternary sets")
    outs = gpu_ctx.graph_build_batch(sets)
    for (syms, ed), adj in zip(sets, outs):
        rp, col = O.graph_csr(syms, ed) if syms.shape[0] else (np.zeros(1, np.uint64), np.zeros(0, np.uint32))
        assert np.array_equal(adj.row_ptr, rp), (syms.shape, ed)
        assert np.array_equal(adj.col, col), (syms.shape, ed)
    bad = [(s.copy(), ed) for s, ed in sets]
    bad[0][0][3, 7] = 3
    with pytest.raises(igc.IGCError) as e:
        gpu_ctx.graph_build_batch(bad)
    assert "binary" in str(e.value)


@pytest.mark.parametrize("n", [6, 10, 16, 20, 32, 40, 64])
def test_forced_predicates_agree(gpu_ctx, n):
    """Landau-Vishkin and Myers forced at the same (n, K), on every word class:
    both give the oracle's CSR, single sets and a batch."""
    rng = np.random.default_rng(6000 + n)
    N = 900
    syms = _binary_set(rng, N, n, N // 3) if n > 32 else rng.integers(0, 4, size=(N, n), dtype=np.uint8)
    if n <= 32:
        _near_copies(rng, syms, N // 3)
    eds = sorted({e for e in (1, 2, 4, 7, 10, 12, 14) if e <= n})
    want = {ed: O.graph_csr(syms, ed) for ed in eds}
    try:
        for pred in (igc.IGC_PRED_LANDAU_VISHKIN, igc.IGC_PRED_MYERS):
            gpu_ctx.set_predicate(pred)
            for ed in eds:
                banner(f"forced predicate {pred} n={n} min_ed={ed}")
                assert gpu_ctx.pair_plan(n, ed, N)["predicate"] == igc.PREDICATES[pred]
                adj = gpu_ctx.graph_build(syms, ed)
                assert np.array_equal(adj.row_ptr, want[ed][0]) and np.array_equal(adj.col, want[ed][1]), (pred, ed)
            outs = gpu_ctx.graph_build_batch([(syms, ed) for ed in eds])
            for ed, adj in zip(eds, outs):
                assert np.array_equal(adj.row_ptr, want[ed][0]) and np.array_equal(adj.col, want[ed][1]), (pred, ed)
    finally:
        gpu_ctx.set_predicate(igc.IGC_PRED_AUTO)


def test_dense_rows_use_every_sort_path(gpu_ctx):
    # random n=12 strings at editDist 5 give rows of ~90 neighbours (wave sort
    # <= 64 and LDS bitonic sort <= 4096); a cluster of 5000 strings sharing 10
    # symbols gives rows above 4096 (bitmap sort)
    rng = np.random.default_rng(8)
    syms = rng.integers(0, 4, size=(12000, 12), dtype=np.uint8)
    syms[:5000, :10] = syms[0, :10]
    banner("dense rows N=12000 n=12 min_ed=5")
    adj = gpu_ctx.graph_build(syms, 5)
    deg = np.diff(adj.row_ptr.astype(np.int64))
    assert deg.max() > 4096 and (deg <= 64).any() and ((deg > 64) & (deg <= 4096)).any()
    rp, col = O.graph_csr(syms, 5)
    assert np.array_equal(adj.row_ptr, rp) and np.array_equal(adj.col, col)


def test_plain_small_config_full(gpu_ctx):
    syms = W.plain_small()
    banner(f"plain_small N={syms.shape[0]} n=11 min_ed=3")
    adj = gpu_ctx.graph_build(syms, 3)
    rp, col = O.graph_csr(syms, 3)
    assert np.array_equal(adj.row_ptr, rp) and np.array_equal(adj.col, col)
    order = igc.greedy_order(adj)
    assert np.array_equal(order.astype(np.int32), O.greedy(rp, col))
    assert igc.verify_dist(syms[order], 3, gpu_ctx)
    assert gpu_ctx.graph_count(syms, 3) == adj.n_edges


def test_published_codebook_edge_counts(gpu_ctx, golden_dir):
    for fn in sorted(os.listdir(golden_dir)):
        if not fn.startswith("published_"):
            continue
        z = np.load(os.path.join(golden_dir, fn))
        syms, ed = z["syms"], int(z["min_ed"])
        banner(f"published {fn} N={syms.shape[0]} min_ed={ed}")
        adj = gpu_ctx.graph_build(syms, ed)
        assert adj.matrix_ones == int(z["ones"])
        check_csr_wellformed(adj)
        # the published codebook is an independent set of this graph
        sel = np.zeros(adj.n_nodes, bool)
        sel[z["order"]] = True
        rows = np.repeat(np.arange(adj.n_nodes), np.diff(adj.row_ptr.astype(np.int64)))
        assert not (sel[rows] & sel[adj.col]).any()


def test_linear_large_config_properties(gpu_ctx):
    syms = W.linear_large()
    banner(f"linear_large N={syms.shape[0]} n=13 min_ed=4")
    adj = gpu_ctx.graph_build(syms, 4)
    check_csr_wellformed(adj)
    # sampled rows against the oracle's row scan
    rng = np.random.default_rng(9)
    for i in rng.integers(0, syms.shape[0], 40):
        assert np.array_equal(adj.neighbors(int(i)), O.row_neighbors(syms, 4, int(i)))
    # repeated builds are identical, and the count-only mode agrees
    adj2 = gpu_ctx.graph_build(syms, 4)
    assert np.array_equal(adj.col, adj2.col) and np.array_equal(adj.row_ptr, adj2.row_ptr)
    assert gpu_ctx.graph_count(syms, 4) == adj.n_edges
    order = igc.greedy_order(adj)
    assert igc.verify_dist(syms[order], 4, gpu_ctx)


def test_batch_matches_single_and_oracle(gpu_ctx):
    rng = np.random.default_rng(21)
    sets = []
    for k in range(40):
        n = int(rng.integers(1, 33))
        N = int(rng.choice([0, 1, 2, 5, 63, 64, 65, 300, 1000, 2500]))
        ed = int(rng.integers(0, n + 3))
        syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        for r in range(0, N // 3):
            src = syms[int(rng.integers(0, N))].copy()
            src[int(rng.integers(0, n))] = int(rng.integers(0, 4))
            syms[r] = src
        sets.append((syms, ed))
    # large thresholds make Myers groups; n > 16 sets make groups on 64-bit
    # words
    sets.append((rng.integers(0, 4, size=(400, 16), dtype=np.uint8), 11))
    sets.append((rng.integers(0, 4, size=(300, 14), dtype=np.uint8), 13))
    sets.append((rng.integers(0, 4, size=(500, 32), dtype=np.uint8), 14))
    sets.append((rng.integers(0, 4, size=(350, 20), dtype=np.uint8), 5))
    banner(f"batch of {len(sets)} uneven sets")
    outs = gpu_ctx.graph_build_batch(sets)
    for (syms, ed), adj in zip(sets, outs):
        rp, col = O.graph_csr(syms, ed) if syms.shape[0] else (np.zeros(1, np.uint64), np.zeros(0, np.uint32))
        assert np.array_equal(adj.row_ptr, rp), (syms.shape, ed)
        assert np.array_equal(adj.col, col), (syms.shape, ed)
    # thread-order flag applies per set, as in the single-set call
    outs4 = gpu_ctx.graph_build_batch(sets[:10], flags=igc.order_threads(4))
    for (syms, ed), adj in zip(sets[:10], outs4):
        if syms.shape[0]:
            rp, col = O.graph_csr(syms, ed, T=4)
            assert np.array_equal(adj.col, col)


def test_batch_prefilter_sets_match_oracle(gpu_ctx):
    """Sets forced through the row filter are sorted and row-filtered inside the batch
    (Landau-Vishkin and Myers groups, on 32- and 64-bit words): the same CSRs as
    the oracle, and as the batch with the filter switched off."""
    rng = np.random.default_rng(64)
    sets = []
    for n, N, ed in [(13, 6000, 4), (16, 4096, 5), (10, 5000, 3), (8, 4500, 2), (11, 5000, 9), (24, 4200, 6),
                     (32, 4100, 9), (12, 300, 3), (14, 4300, 15)]:
        syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        _near_copies(rng, syms, N // 3)
        syms[N // 2: N // 2 + 20] = 1  # one-symbol strings: all counts in one byte
        sets.append((syms, ed))
    banner(f"prefiltered batch of {len(sets)} sets")
    gpu_ctx.set_prefilter(igc.IGC_PREFILTER_ALWAYS)
    try:
        outs = gpu_ctx.graph_build_batch(sets)
    finally:
        gpu_ctx.set_prefilter(True)
    for (syms, ed), adj in zip(sets, outs):
        rp, col = O.graph_csr(syms, ed)
        assert np.array_equal(adj.row_ptr, rp), (syms.shape, ed)
        assert np.array_equal(adj.col, col), (syms.shape, ed)
    gpu_ctx.set_prefilter(False)
    try:
        outs_off = gpu_ctx.graph_build_batch(sets)
    finally:
        gpu_ctx.set_prefilter(True)
    for a, b in zip(outs, outs_off):
        assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col, b.col)


def test_batch_device_sets_match_host_batch(gpu_ctx):
    """Sets in caller device buffers (gathered into one symbol array by a single
    launch): the union CSR's row slices equal the host-input batch's per-set
    CSRs; empty sets and byte counts that are not multiples of 16 included."""
    import ctypes
    igc.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")  # the HIP runtime libigc.so is linked against
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    rng = np.random.default_rng(77)
    sets = []
    for N, n, ed in [(0, 11, 3), (1, 7, 2), (2, 13, 4), (333, 11, 3), (0, 5, 1), (1000, 13, 4), (17, 3, 1),
                     (2500, 16, 5), (4097, 10, 3), (5, 1, 1), (700, 24, 6), (300, 40, 6)]:
        syms = rng.integers(0, 2 if n > 32 else 4, size=(N, n), dtype=np.uint8)
        if n > 32:  # binary words: near copies by bit flips
            for r in range(N // 3):
                syms[r] = syms[int(rng.integers(0, N))]
                syms[r, int(rng.integers(0, n))] ^= 1
        elif N:
            _near_copies(rng, syms, N // 3)
        sets.append((syms, ed))
    banner(f"device batch of {len(sets)} sets")
    bufs = []
    for syms, _ in sets:
        p = ctypes.c_void_p()
        if syms.size:
            a = np.ascontiguousarray(syms)
            assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        bufs.append(p.value or 0)
    spec = [(p, s.shape[0], s.shape[1], ed) for p, (s, ed) in zip(bufs, sets)]
    try:
        _check_device_batch(gpu_ctx, sets, spec)
    finally:
        for p in bufs:
            if p:
                hip.hipFree(ctypes.c_void_p(p))


def _check_device_batch(gpu_ctx, sets, spec):
    for _ in range(2):  # the second call reuses the context's buffers
        d, off, edges = gpu_ctx.graph_build_batch_device(spec)
        union = gpu_ctx.device_csr_to_host(d)
        outs = gpu_ctx.graph_build_batch(sets)
        for k, ((syms, ed), adj) in enumerate(zip(sets, outs)):
            a, b = int(off[k]), int(off[k + 1])
            assert b - a == syms.shape[0]
            rp = union.row_ptr[a:b + 1] - union.row_ptr[a]
            col = union.col[int(union.row_ptr[a]):int(union.row_ptr[b])] - a
            assert np.array_equal(rp, adj.row_ptr), (k, syms.shape, ed)
            assert np.array_equal(col, adj.col), (k, syms.shape, ed)
            assert int(edges[k]) == adj.n_edges


def _upper_of(rp, col):
    """Upper-triangle CSR (each row keeps its neighbours above the row) of a
    symmetric one."""
    N = rp.size - 1
    rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(rp.astype(np.int64)))
    keep = col.astype(np.int64) > rows
    urp = np.zeros(N + 1, dtype=np.uint64)
    urp[1:] = np.cumsum(np.bincount(rows[keep], minlength=N))
    return urp, col[keep]


def test_multi_chunk_host_copies(gpu_ctx):
    """Host CSRs spanning several 4 MiB staging chunks (moved by worker threads,
    and in a batch split among the sets a chunk holds): single-set and batched
    results equal the oracle, with the thread order and the upper triangle
    applied after the streamed copy."""
    rng = np.random.default_rng(5)
    big = rng.integers(0, 4, size=(16000, 11), dtype=np.uint8)  # ~4.4e6 entries, ~17 MiB of col
    banner("multi-chunk host copies: N=16000 n=11 ed=5")
    rp, col = O.graph_csr(big, 5)
    assert col.size * 4 > 3 * (4 << 20)
    adj = gpu_ctx.graph_build(big, 5)
    assert np.array_equal(adj.row_ptr, rp) and np.array_equal(adj.col, col)
    small = [rng.integers(0, 4, size=(N, n), dtype=np.uint8) for N, n in ((900, 10), (1, 12), (2500, 13))]
    sets = [(small[0], 3), (big, 5), (small[1], 2), (small[2], 4), (big[:9000].copy(), 5)]
    singles = [gpu_ctx.graph_build(s, ed) for s, ed in sets]
    outs = gpu_ctx.graph_build_batch(sets)
    for k, (a, b) in enumerate(zip(outs, singles)):
        assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col, b.col), k
    assert np.array_equal(outs[1].col, col)
    rp4, col4 = O.graph_csr(big, 5, T=4)
    outs4 = gpu_ctx.graph_build_batch(sets, flags=igc.order_threads(4))
    assert np.array_equal(outs4[1].row_ptr, rp4) and np.array_equal(outs4[1].col, col4)
    urp, ucol = _upper_of(rp, col)
    up = gpu_ctx.graph_build_batch(sets, flags=igc.IGC_UPPER_TRIANGLE)
    assert np.array_equal(up[1].row_ptr, urp) and np.array_equal(up[1].col, ucol)
    for (s, _), o, full in zip(sets, up, singles):
        u_rp, u_col = _upper_of(full.row_ptr, full.col)
        assert np.array_equal(o.row_ptr, u_rp) and np.array_equal(o.col, u_col)


def test_batched_config_sets_alone_and_in_the_mix(gpu_ctx):
    # a scaled-down batched workload: each set's graph is the same alone and in the batch
    sets = W.batched(seed=3, sets=24, scale=0.05)
    banner(f"batched workload {len(sets)} sets, pairs={sum(s.pairs for s in sets):.3e}")
    outs = gpu_ctx.graph_build_batch([(s.syms, s.edit_dist) for s in sets])
    for s, adj in zip(sets, outs):
        alone = gpu_ctx.graph_build(s.syms, s.edit_dist)
        assert np.array_equal(adj.row_ptr, alone.row_ptr) and np.array_equal(adj.col, alone.col)
    for k in range(4):
        rp, col = O.graph_csr(sets[k].syms, sets[k].edit_dist)
        assert np.array_equal(outs[k].col, col)


def test_kmeans_matches_oracle_bit_exact(gpu_ctx):
    rng = np.random.default_rng(31)
    for N, n, k, iters, seed in [(20000, 16, 64, 10, 1), (5000, 11, 8, 3, 2), (30000, 13, 256, 10, 3), (777, 5, 7, 4, 9),
                                 (8000, 24, 32, 5, 4), (3000, 32, 100, 3, 5), (2000, 17, 9, 4, 6)]:
        banner(f"kmeans N={N} n={n} k={k} iters={iters}")
        syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        a_o, C_o = O.kmeans(syms, k, iters, seed)
        a_g, C_g = gpu_ctx.kmeans(syms, k, iters, seed)
        assert np.array_equal(a_g, a_o)
        assert np.array_equal(C_g.view(np.uint32), C_o.view(np.uint32))  # bitwise FP32


def test_kmeans_graph_replays_equal_eager(gpu_ctx):
    # every replay count of the captured Lloyd iteration against eager launches,
    # with eager work and synchronisation between the graph runs
    rng = np.random.default_rng(32)
    syms = rng.integers(0, 4, size=(12000, 16), dtype=np.uint8)
    for iters in (1, 2, 3, 5, 8):
        banner(f"kmeans graph vs eager iters={iters}")
        a_e, C_e = gpu_ctx.kmeans(syms, 32, iters, 5, eager=True)
        gpu_ctx.graph_build(syms[:500], 4)  # unrelated work between replays
        a_g, C_g = gpu_ctx.kmeans(syms, 32, iters, 5, eager=False)
        assert np.array_equal(a_e, a_g) and np.array_equal(C_e, C_g)
        a_o, _ = O.kmeans(syms, 32, iters, 5)
        assert np.array_equal(a_g, a_o)


def test_assign_fp32_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(62)
    for N, n, k in [(5000, 13, 37), (3000, 24, 9), (700, 5, 60), (20000, 16, 256)]:
        banner(f"assign_fp32 N={N} n={n} k={k}")
        syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        C = rng.random((k, 4 * n), dtype=np.float32)
        for c in range(min(k, 5)):  # one-hot copies of real strings
            C[c] = 0.0
            C[c, 4 * np.arange(n) + syms[c]] = 1.0
        assert np.array_equal(gpu_ctx.assign_fp32(syms, C), O.assign_fp32(syms, C))
    # the K-means result's own centroids assign like its final step
    syms = rng.integers(0, 4, size=(8000, 12), dtype=np.uint8)
    a, C = gpu_ctx.kmeans(syms, 16, 4, 7)
    assert np.array_equal(gpu_ctx.assign_fp32(syms, C), a)


def _forms(gpu_ctx, build):
    out = {}
    for form in (igc.IGC_CSR_WIDE_KEYS, igc.IGC_CSR_AUTO):
        gpu_ctx.set_csr_form(form)
        try:
            out[form] = build()
        finally:
            gpu_ctx.set_csr_form(igc.IGC_CSR_AUTO)
    return out


def test_csr_forms_agree(gpu_ctx):
    """The segmented 32-bit CSR form (split pass on the row's high bits) and the
    64-bit-key form give the same CSR: single sets from 2 to 140,000 strings
    (no split up to 65,536 nodes, a 2-bit split at 70,000, 4-bit at 140,000),
    dense and sparse, rows of a few entries up to all 17,499 other strings,
    the upper-triangle form, and a batch whose largest set has 100,000
    strings (neighbours stored relative to their set's first node, 2-bit
    split)."""
    rng = np.random.default_rng(66)
    cases = [(2, 5, 3), (700, 9, 4), (6000, 13, 5), (20000, 10, 4), (70000, 11, 3), (140000, 12, 3)]
    for N, n, ed in cases:
        banner(f"csr forms N={N} n={n} ed={ed}")
        syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        _near_copies(rng, syms, N // 4)
        out = _forms(gpu_ctx, lambda: (gpu_ctx.graph_build(syms, ed), gpu_ctx.graph_build(syms, ed,
                                                                                          igc.IGC_UPPER_TRIANGLE)))
        (a, au) = out[igc.IGC_CSR_WIDE_KEYS]
        for form in (igc.IGC_CSR_AUTO,):
            b, bu = out[form]
            assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col, b.col), (N, form)
            assert np.array_equal(au.row_ptr, bu.row_ptr) and np.array_equal(au.col, bu.col), (N, form)
        b = out[igc.IGC_CSR_AUTO][0]
        check_csr_wellformed(b)
        if N <= 6000:
            rp, col = O.graph_csr(syms, ed)
            assert np.array_equal(b.row_ptr, rp) and np.array_equal(b.col, col)
    # long rows: clusters of near-identical strings (rows of a few thousand
    # entries), and one cluster where every pair conflicts
    for per, clusters in ((3000, 6), (6000, 2), (17500, 1)):
        base = rng.integers(0, 4, size=(clusters, 16), dtype=np.uint8)
        syms = np.repeat(base, per, axis=0)
        pos = rng.integers(0, 16, size=syms.shape[0])
        syms[np.arange(syms.shape[0]), pos] = rng.integers(0, 4, size=syms.shape[0], dtype=np.uint8)
        syms = syms[rng.permutation(syms.shape[0])]
        banner(f"csr forms long rows: {clusters} x {per}")
        out = _forms(gpu_ctx, lambda: gpu_ctx.graph_build(syms, 4))
        a = out[igc.IGC_CSR_WIDE_KEYS]
        assert np.diff(a.row_ptr.astype(np.int64)).max() >= per - 1
        for form in (igc.IGC_CSR_AUTO,):
            assert np.array_equal(a.row_ptr, out[form].row_ptr) and np.array_equal(a.col, out[form].col), form
    sets = W.batched(seed=5, sets=12, scale=0.1)
    big = rng.integers(0, 4, size=(100000, 12), dtype=np.uint8)
    _near_copies(rng, big, 20000)
    data = [(s.syms, s.edit_dist) for s in sets[:6]] + [(big, 4)] + [(s.syms, s.edit_dist) for s in sets[6:]]
    banner(f"csr forms batch of {len(data)}")
    out = _forms(gpu_ctx, lambda: gpu_ctx.graph_build_batch(data))
    for form in (igc.IGC_CSR_AUTO,):
        for a, b in zip(out[igc.IGC_CSR_WIDE_KEYS], out[form]):
            assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col, b.col)
    with pytest.raises(igc.IGCError):
        gpu_ctx.set_csr_form(2)


def test_upper_triangle_csr(gpu_ctx):
    rng = np.random.default_rng(63)
    for N, n, ed in [(700, 9, 4), (6000, 13, 5), (5000, 20, 6)]:  # 6000: with the row prefilter
        banner(f"upper triangle N={N} n={n} ed={ed}")
        syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        syms[: N // 4] = syms[N // 4: 2 * (N // 4)]
        full = gpu_ctx.graph_build(syms, ed)
        gpu_ctx.set_prefilter(igc.IGC_PREFILTER_ALWAYS if N == 6000 else igc.IGC_PREFILTER_AUTO)
        try:
            up = gpu_ctx.graph_build(syms, ed, igc.IGC_UPPER_TRIANGLE)
        finally:
            gpu_ctx.set_prefilter(True)
        assert up.n_edges == full.n_edges and up.col.size == full.n_edges
        for x in range(0, N, max(1, N // 200)):
            row = full.col[full.row_ptr[x]:full.row_ptr[x + 1]]
            assert np.array_equal(up.col[up.row_ptr[x]:up.row_ptr[x + 1]], row[row > x]), x
    sets = [(rng.integers(0, 4, size=(400, 10), dtype=np.uint8), 3), (rng.integers(0, 4, size=(300, 12), dtype=np.uint8), 4)]
    outs = gpu_ctx.graph_build_batch(sets, flags=igc.IGC_UPPER_TRIANGLE)
    for (s, ed), o in zip(sets, outs):
        full = gpu_ctx.graph_build(s, ed)
        assert o.col.size == full.n_edges
        for x in range(s.shape[0]):
            row = full.col[full.row_ptr[x]:full.row_ptr[x + 1]]
            assert np.array_equal(o.col[o.row_ptr[x]:o.row_ptr[x + 1]], row[row > x])


def refine_oracle(syms, ed, k, iters, rounds, conv, seed):
    """The clustered loop (src/SparseMat.cpp:918-1095) with the oracle's pieces."""
    cur = np.arange(syms.shape[0])
    sizes = []
    for r in range(rounds):
        M = cur.size
        keff = k
        if M // keff > 200000:
            keff = max(2, (M + 199999) // 200000)
        keff = min(keff, M)
        a, _ = O.kmeans(syms[cur], keff, iters, seed + 0x9E37 * r)
        nxt = []
        for c in range(keff):
            mem = np.flatnonzero(a == c)
            if mem.size == 0:
                continue
            rp, col = O.graph_csr(syms[cur[mem]], ed)
            nxt.extend(cur[mem[O.greedy(rp, col)]].tolist())
        sizes.append(len(nxt))
        cur = np.array(nxt, dtype=np.int64)
        if len(sizes) >= conv and all(sizes[-1] == sizes[-1 - q] for q in range(1, conv)):
            break
    return cur


def test_refine_matches_oracle(gpu_ctx):
    syms = W.vt_refine(seed=4, total=6000)
    banner(f"refine N={syms.shape[0]} n=16 ed=5 k=8")
    got, stats = gpu_ctx.refine(syms, 5, 8, kmeans_iters=5, max_rounds=3, converge_n=2, seed=7, host_threads=4)
    want = refine_oracle(syms, 5, 8, 5, 3, 2, 7)
    assert np.array_equal(got.astype(np.int64), want)
    assert stats["rounds"] and stats["rounds"][0]["candidates"] == syms.shape[0]
    # the codebook of the last round is independent inside every cluster; across
    # clusters the reference's loop guarantees nothing, so only the size is checked
    assert 0 < got.size <= syms.shape[0]


def test_batched_full_config_properties(gpu_ctx):
    """The full uneven-sets workload (64 sets, n 10..16, editDist 3..5, up to
    150,000 strings, 8.2e10 pairs): every set's edge count in the batch equals its
    single build, sampled sets are identical alone and in the mix, sampled rows
    equal the oracle's row scan, and the device-input batch gives the same union."""
    sets = W.batched(seed=3, sets=64, scale=1.0)
    banner(f"full batched workload: {len(sets)} sets, {sum(s.pairs for s in sets):.3e} pairs")
    outs = gpu_ctx.graph_build_batch([(s.syms, s.edit_dist) for s in sets])
    rng = np.random.default_rng(12)
    sampled = set(int(k) for k in rng.choice(len(sets), 6, replace=False))
    for k, (s, o) in enumerate(zip(sets, outs)):
        assert o.n_nodes == s.syms.shape[0] and int(o.row_ptr[-1]) == 2 * o.n_edges
        if k in sampled:
            alone = gpu_ctx.graph_build(s.syms, s.edit_dist)
            assert np.array_equal(o.row_ptr, alone.row_ptr) and np.array_equal(o.col, alone.col), k
            for i in rng.integers(0, s.syms.shape[0], 3):
                assert np.array_equal(o.neighbors(int(i)), O.row_neighbors(s.syms, s.edit_dist, int(i))), (k, i)
        else:
            assert gpu_ctx.graph_count(s.syms, s.edit_dist) == o.n_edges, k
    total = sum(o.n_edges for o in outs)
    del outs
    counts_only = sum(gpu_ctx.graph_count(s.syms, s.edit_dist) for s in sets)
    assert counts_only == total


def test_refine_full_config_deterministic(gpu_ctx):
    """The full cluster-refinement workload (500,000 strings n=16 with the
    syndrome mod 17, editDist 5, k=256, 10 Lloyd iterations, 5 rounds) twice:
    the same codebook, and each round's candidates are the previous round's
    codewords."""
    syms = W.vt_refine(seed=4)
    banner(f"full refinement N={syms.shape[0]}")
    a, sa = gpu_ctx.refine(syms, 5, 256, kmeans_iters=10, max_rounds=5, converge_n=3, seed=1, host_threads=8)
    b, sb = gpu_ctx.refine(syms, 5, 256, kmeans_iters=10, max_rounds=5, converge_n=3, seed=1, host_threads=8)
    assert np.array_equal(a, b)
    rounds = sa["rounds"]
    assert rounds[0]["candidates"] == syms.shape[0]
    for r0, r1 in zip(rounds, rounds[1:]):
        assert r1["candidates"] == r0["codewords"] and r0["codewords"] <= r0["candidates"]
    assert a.size == rounds[-1]["codewords"] and len(np.unique(a)) == a.size
    assert [r["codewords"] for r in rounds] == [r["codewords"] for r in sb["rounds"]]


def test_batch_symbols_over_several_staging_chunks(gpu_ctx):
    """A host batch whose concatenated symbols (~30 MB) fill several 8 MiB
    pinned staging chunks, set boundaries falling inside chunks: every set's CSR
    equals its single build (which copies its symbols by itself)."""
    rng = np.random.default_rng(17)
    sets = []
    for N, n, ed in ((700_000, 11, 2), (3, 11, 2), (900_000, 11, 2), (1, 9, 1), (1_100_000, 12, 2), (999, 13, 3)):
        syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        if N > 10:
            syms[: N // 50] = syms[N // 50: 2 * (N // 50)]  # duplicates: edges at editDist 2
        sets.append((syms, ed))
    total = sum(s.nbytes for s, _ in sets)
    banner(f"host batch with {total / 2**20:.1f} MiB of symbols")
    assert total > 3 * (8 << 20)
    outs = gpu_ctx.graph_build_batch(sets)
    for k, ((s, ed), o) in enumerate(zip(sets, outs)):
        single = gpu_ctx.graph_build(s, ed)
        assert np.array_equal(o.row_ptr, single.row_ptr) and np.array_equal(o.col, single.col), k
        if s.shape[0] > 10:
            assert o.n_edges >= s.shape[0] // 50


def test_concurrent_calls_from_host_threads(gpu_ctx):
    """The C ABI is callable from several host threads at once, on one context
    (calls serialise on its lock) and on separate contexts (their work overlaps):
    every result equals the sequential one, with host CSRs large enough to use the
    threaded staging copy and the process-wide host block cache."""
    import threading
    rng = np.random.default_rng(91)
    jobs = [(rng.integers(0, 4, size=(N, n), dtype=np.uint8), ed)
            for N, n, ed in ((12000, 11, 5), (3000, 13, 4), (9000, 10, 4), (500, 16, 5), (7000, 12, 5))]
    ref = [gpu_ctx.graph_build(s, ed) for s, ed in jobs]
    banner(f"{len(jobs)} jobs from 4 threads, shared and own contexts")
    own = [igc.Context(0) for _ in range(2)]
    errors, results = [], {}

    def worker(w):
        try:
            ctx = gpu_ctx if w < 2 else own[w - 2]
            for rep in range(3):
                for k, (s, ed) in enumerate(jobs):
                    a = ctx.graph_build(s, ed)
                    results[(w, rep, k)] = (a.row_ptr.copy(), a.col.copy())
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    for c in own:
        c.close()
    assert not errors, errors
    assert len(results) == 4 * 3 * len(jobs)
    for (w, rep, k), (rp, col) in results.items():
        assert np.array_equal(rp, ref[k].row_ptr) and np.array_equal(col, ref[k].col), (w, rep, k)


def test_server_results_same_alone_and_in_the_mix(gpu_ctx):
    specs = W.serving(seed=5, jobs=40, mu=np.log(3000), sigma=1.0, lo=100, hi=20000)
    data = [(sp.draw(), sp.edit_dist) for sp in specs]
    banner(f"server: {len(data)} jobs, pairs={sum(sp.pairs for sp in specs):.3e}")
    srv = igc.Server(0, pipelines=2, host_threads=4)
    try:
        ids = [srv.submit(s, ed) for s, ed in data]          # all at once: they share launches
        mixed = [srv.wait(j) for j in ids]
        assert max(r["batch_jobs"] for r in mixed) > 1        # work really was batched
        alone = []
        for s, ed in data[:12]:                               # one at a time: a launch each
            r = srv.wait(srv.submit(s, ed))
            assert r["batch_jobs"] == 1
            alone.append(r)
    finally:
        srv.close()
    for (s, ed), r in zip(data, mixed):
        assert r["status"] == 0
        adj = gpu_ctx.graph_build(s, ed)                      # the single-set path
        assert r["n_edges"] == adj.n_edges
        assert np.array_equal(r["order"], igc.greedy_order(adj))
        assert r["submit_s"] <= r["admit_s"] <= r["gpu_done_s"] <= r["done_s"]
    for r_alone, r_mix in zip(alone, mixed):
        assert np.array_equal(r_alone["order"], r_mix["order"])
    # one job checked all the way down to the oracle
    s, ed = data[0]
    rp, col = O.graph_csr(s, ed)
    assert np.array_equal(mixed[0]["order"].astype(np.int32), O.greedy(rp, col))


def test_server_admission_budget(gpu_ctx):
    rng = np.random.default_rng(41)
    big = W.filtered_uniform(rng, 20000, 12)
    srv = igc.Server(0, hbm_budget_bytes=6 << 20, pipelines=1, host_threads=2)   # 6 MiB
    try:
        r = srv.wait(srv.submit(big, 4))                      # can never fit: refused
        assert r["status"] == igc.IGC_ERR_ADMISSION
        small = [W.filtered_uniform(rng, 300, 12) for _ in range(6)]
        ids = [srv.submit(s, 4) for s in small]               # fit only a few at a time
        res = [srv.wait(j) for j in ids]
        assert all(x["status"] == 0 for x in res)
        for s, x in zip(small, res):
            assert np.array_equal(x["order"], igc.greedy_order(gpu_ctx.graph_build(s, 4)))
    finally:
        srv.close()


def test_autotuned_shapes_give_identical_graphs(gpu_ctx):
    rng = np.random.default_rng(51)
    syms = W.filtered_uniform(rng, 9000, 12)
    rp, col = O.graph_csr(syms, 4)
    ents = gpu_ctx.autotune(12, 4, syms.shape[0], reps=2)
    assert len(ents) == 12 and all(e["ms"] > 0 for e in ents)
    fastest = min(e["ms"] for e in ents)
    pick = next(e for e in ents if e["ms"] <= fastest * np.float32(1.02))  # first within 2%, in trial order
    assert gpu_ctx.tuned_shape(12, 4, syms.shape[0]) == (pick["cpl"], pick["q"])
    plan = gpu_ctx.pair_plan(12, 4, syms.shape[0])
    assert plan["predicate"] == "landau_vishkin" and plan["k"] == 3 and plan["lane_cols"] == 2
    assert (plan["slots"], plan["q"]) == (pick["cpl"], pick["q"])
    assert plan["tile_cols"] == 256 * plan["slots"] * 2 and plan["tile_rows"] == plan["tile_cols"] // plan["q"]
    assert not plan["row_filter"]
    big = gpu_ctx.pair_plan(12, 4, 200000)  # filtered: one register of columns per lane
    assert big["row_filter"] and big["slots"] == 1
    assert gpu_ctx.pair_plan(24, 20, 100)["predicate"] == "myers"
    assert gpu_ctx.pair_plan(24, 40, 100)["predicate"] == "all_pairs"
    adj = gpu_ctx.graph_build(syms, 4)  # built with the tuned shape
    assert np.array_equal(adj.row_ptr, rp) and np.array_equal(adj.col, col)
    # 64-bit words (Landau-Vishkin at n=24, K=5)
    syms = rng.integers(0, 4, size=(5000, 24), dtype=np.uint8)
    rp, col = O.graph_csr(syms, 6)
    ents = gpu_ctx.autotune(24, 6, syms.shape[0], reps=1)
    assert len(ents) == 12
    adj = gpu_ctx.graph_build(syms, 6)
    assert np.array_equal(adj.row_ptr, rp) and np.array_equal(adj.col, col)


def test_cli_writes_the_reference_codebook_file(gpu_ctx, tmp_path):
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    syms = W.plain_small(seed=3, count=2500)
    cf = tmp_path / "cands.txt"
    cf.write_text("\n".join(igc.to_strings(syms)) + "\n")
    out = subprocess.run([os.path.join(root, "indexgen_amd", "igc_cli"), "--cands", str(cf), "--min-ed", "3",
                          "--out", str(tmp_path), "--verify", "--order-threads", "4"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "Vector distance SUCCESS" in out.stdout
    rp, col = O.graph_csr(syms, 3, T=4)
    want = igc.to_strings(syms[O.greedy(rp, col)])
    files = [f for f in os.listdir(tmp_path) if f.startswith("CodeSize-")]
    assert files == [f"CodeSize-{len(want):07d}_CodeLen-11_MinED-3.txt"]
    body = open(tmp_path / files[0]).read().split("=========================================== \n")[1].split()
    assert body == want
    # --graph-only: what the reference's per-call GPU script does (pair file, no codebook)
    pf = tmp_path / "edges.bin"
    out = subprocess.run([os.path.join(root, "indexgen_amd", "igc_cli"), "--cands", str(cf), "--min-ed", "3",
                          "--pair-file", str(pf), "--graph-only"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    pairs = np.fromfile(pf, dtype="<i4").reshape(-1, 2)
    assert (pairs[:, 0] < pairs[:, 1]).all()
    rp1, col1 = O.graph_csr(syms, 3)
    rows = np.repeat(np.arange(syms.shape[0]), np.diff(rp1.astype(np.int64)))
    keep = col1.astype(np.int64) > rows
    want_pairs = np.stack([rows[keep], col1[keep].astype(np.int64)], 1)
    got = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))].astype(np.int64)
    assert np.array_equal(got, want_pairs)


def test_cli_refine_writes_the_refined_codebook(gpu_ctx, tmp_path):
    """igc_cli --refine runs the clustered loop (igc_refine, seed 1) and writes its
    codebook file: the words equal Context.refine's for the same parameters."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    syms = W.vt_refine(seed=4, total=5000)
    cf = tmp_path / "cands.txt"
    cf.write_text("\n".join(igc.to_strings(syms)) + "\n")
    out = subprocess.run([os.path.join(root, "indexgen_amd", "igc_cli"), "--cands", str(cf), "--min-ed", "5",
                          "--out", str(tmp_path), "--refine", "8", "--kmeans-iters", "5", "--rounds", "3",
                          "--converge", "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    order, _ = gpu_ctx.refine(syms, 5, 8, kmeans_iters=5, max_rounds=3, converge_n=2, seed=1, host_threads=4)
    want = igc.to_strings(syms[order])
    files = [f for f in os.listdir(tmp_path) if f.startswith("CodeSize-")]
    assert files == [f"CodeSize-{len(want):07d}_CodeLen-16_MinED-5.txt"]
    body = open(tmp_path / files[0]).read().split("=========================================== \n")[1].split()
    assert body == want
