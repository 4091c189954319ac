"""The continuous-batching server and the build's memory paths on the GPU.

- The whole seeded serving stream (2,000 jobs, n in 11..16, editDist in 3..5,
  sizes lognormal(ln 30000, 1.2) clipped to [500, 400000]) submitted at once:
  every job succeeds, its edge count equals the count-only build, sampled jobs
  (the five largest among them) give the codebook of their single build, and
  the smallest ones the oracle's.
- A job with bad symbols is refused at submission and cannot fail the jobs it
  would have shared a batch with.
- A job with far more edges than its admission estimate either completes with
  the single build's codebook or fails cleanly with IGC_ERR_OOM; either way the
  server's reserved HBM never exceeds its budget, and it keeps serving.
- A first edge buffer forced far too small is regrown to the exact count and
  the pairs replayed: the same graph, single set and batch.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import indexgen_amd as igc
from indexgen_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _draw_all(specs):
    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda sp: sp.draw(), specs))


def test_full_serving_stream(gpu_ctx):
    """On four pipelines (the most a server runs)."""
    specs = W.serving(seed=5)
    assert len(specs) == 2000
    data = _draw_all(specs)
    print(f"[server] drew {sum(d.shape[0] for d in data)} strings for {len(specs)} jobs", flush=True)
    srv = igc.Server(0, pipelines=4, host_threads=16)
    try:
        ids = [srv.submit(d, sp.edit_dist) for d, sp in zip(data, specs)]
        res = [srv.wait(j) for j in ids]
        info = srv.stats()
    finally:
        srv.close()
    bad = [(k, r["status"], r["error"]) for k, r in enumerate(res) if r["status"] != 0]
    assert bad == []
    assert info["max_in_use_bytes"] <= 200 << 30
    print(f"[server] 2000 jobs done, {info['batches']} batches, edges {sum(r['n_edges'] for r in res)}",
          flush=True)
    sizes = np.array([sp.size for sp in specs])
    rng = np.random.default_rng(12)
    sample = sorted(set(np.argsort(sizes)[-5:].tolist()) | set(rng.choice(len(specs), 15, replace=False).tolist()))
    for k in range(len(specs)):  # every job: edge count of the count-only build
        assert res[k]["n_edges"] == gpu_ctx.graph_count(data[k], specs[k].edit_dist), k
        assert res[k]["n_nodes"] == sizes[k]
    def single(k):  # one build + greedy; the host CSR is dropped before the next
        return np.array_equal(res[k]["order"], igc.greedy_order(gpu_ctx.graph_build(data[k], specs[k].edit_dist)))

    with ThreadPoolExecutor(4) as ex:
        same = list(ex.map(single, sample))
    assert [k for k, ok in zip(sample, same) if not ok] == []
    for k in np.argsort(sizes)[:3]:  # the smallest jobs against the oracle
        rp, col = O.graph_csr(data[k], specs[k].edit_dist)
        assert np.array_equal(res[k]["order"].astype(np.int32), O.greedy(rp, col)), k


def test_bad_job_refused_at_submission(gpu_ctx):
    rng = np.random.default_rng(3)
    good = [W.filtered_uniform(rng, 700, 12) for _ in range(6)]
    bad = good[2].copy()
    bad[100, 3] = 4
    binary_bad = rng.integers(0, 2, size=(300, 40), dtype=np.uint8)
    binary_bad[7, 7] = 2
    srv = igc.Server(0, pipelines=1, host_threads=2)
    try:
        ids = [srv.submit(g, 4) for g in good[:3]]
        with pytest.raises(igc.IGCError) as e:
            srv.submit(bad, 4)
        assert e.value.code == igc.IGC_ERR_INVALID and "above 3" in str(e.value)
        with pytest.raises(igc.IGCError) as e:
            srv.submit(binary_bad, 5)
        assert e.value.code == igc.IGC_ERR_INVALID and "above 1" in str(e.value)
        ids += [srv.submit(g, 4) for g in good[3:]]
        res = [srv.wait(j) for j in ids]
    finally:
        srv.close()
    for g, r in zip(good, res):
        assert r["status"] == 0 and r["error"] == ""
        assert np.array_equal(r["order"], igc.greedy_order(gpu_ctx.graph_build(g, 4)))


def _dense_set(rng, bases=12, per=1000, n=16):
    """Strings within two substitutions of a few bases: ~6M edges at editDist 5,
    far above the density estimate of random length-16 strings (~1.1M)."""
    out = []
    for _ in range(bases):
        b = rng.integers(0, 4, size=n, dtype=np.uint8)
        v = np.repeat(b[None, :], per, axis=0)
        for _ in range(2):
            pos = rng.integers(0, n, size=per)
            v[np.arange(per), pos] = rng.integers(0, 4, size=per, dtype=np.uint8)
        out.append(v)
    return np.concatenate(out)


def test_underestimated_job_within_budget(gpu_ctx):
    rng = np.random.default_rng(8)
    dense = _dense_set(rng)
    small = W.filtered_uniform(rng, 500, 12)
    want = igc.greedy_order(gpu_ctx.graph_build(dense, 5))
    E = gpu_ctx.graph_count(dense, 5)
    assert E > 5_000_000
    # a budget the estimate (~52 MB) fits but the real graph (edge list and
    # CSR, ~100 MB) does not: a clean IGC_ERR_OOM
    budget = 64 << 20
    srv = igc.Server(0, hbm_budget_bytes=budget, pipelines=2, host_threads=2)
    try:
        r = srv.wait(srv.submit(dense, 5))
        assert r["status"] == igc.IGC_ERR_OOM, r
        assert "budget" in r["error"]
        r2 = srv.wait(srv.submit(small, 4))  # still serving
        assert r2["status"] == 0
        assert np.array_equal(r2["order"], igc.greedy_order(gpu_ctx.graph_build(small, 4)))
        info = srv.stats()
        assert info["max_in_use_bytes"] <= budget
    finally:
        srv.close()
    # a budget the real graph fits: it completes, in a batch with other jobs
    budget = 2 << 30
    srv = igc.Server(0, hbm_budget_bytes=budget, pipelines=2, host_threads=2)
    try:
        ids = [srv.submit(small, 4), srv.submit(dense, 5), srv.submit(small, 3)]
        res = [srv.wait(j) for j in ids]
        info = srv.stats()
    finally:
        srv.close()
    assert [x["status"] for x in res] == [0, 0, 0]
    assert res[1]["n_edges"] == E
    assert np.array_equal(res[1]["order"], want)
    assert info["max_in_use_bytes"] <= budget


def test_edge_buffer_regrow_and_replay(gpu_ctx):
    """A forced tiny first edge buffer (igc_ctx_set_edge_cap) overflows; the build
    regrows it to the exact count and replays the pairs.  Single and batched
    builds, quaternary n = 10..16, n = 20 (no density prior) and binary n = 40,
    all equal to the oracle; with the cap cleared no retry happens."""
    rng = np.random.default_rng(21)
    syms = W.filtered_uniform(rng, 6000, 11)
    rp, col = O.graph_csr(syms, 4)
    assert col.size // 2 > 10000
    try:
        gpu_ctx.set_edge_cap(1000)
        adj = gpu_ctx.graph_build(syms, 4)
        assert gpu_ctx.last_timing().retries >= 1
        assert np.array_equal(adj.row_ptr, rp) and np.array_equal(adj.col, col)
        rp4, col4 = O.graph_csr(syms, 4, T=4)
        adj = gpu_ctx.graph_build(syms, 4, igc.order_threads(4))
        assert np.array_equal(adj.row_ptr, rp4) and np.array_equal(adj.col, col4)
        assert gpu_ctx.graph_count(syms, 4) == col.size // 2
        sets = [(W.filtered_uniform(rng, N, n), ed) for N, n, ed in ((3000, 12, 5), (500, 10, 3), (2500, 16, 6))]
        outs = gpu_ctx.graph_build_batch(sets)
        assert gpu_ctx.last_timing().retries >= 1
        for (s, ed), a in zip(sets, outs):
            r, c = O.graph_csr(s, ed)
            assert np.array_equal(a.row_ptr, r) and np.array_equal(a.col, c)
        # lengths without a density prior (n > 16: 64-bit words; n > 32: binary)
        wide = [(W.filtered_uniform(rng, 2500, 20), 9),
                (W.dedup_first(rng.integers(0, 2, size=(2500, 40), dtype=np.uint8)), 9)]
        for s, ed in wide:
            r, c = O.graph_csr(s, ed)
            assert c.size // 2 > 1000
            adj = gpu_ctx.graph_build(s, ed)
            assert gpu_ctx.last_timing().retries >= 1
            assert np.array_equal(adj.row_ptr, r) and np.array_equal(adj.col, c)
        outs = gpu_ctx.graph_build_batch(wide)
        assert gpu_ctx.last_timing().retries >= 1
        for (s, ed), a in zip(wide, outs):
            r, c = O.graph_csr(s, ed)
            assert np.array_equal(a.row_ptr, r) and np.array_equal(a.col, c)
    finally:
        gpu_ctx.set_edge_cap(0)
    adj = gpu_ctx.graph_build(syms, 4)
    assert gpu_ctx.last_timing().retries == 0
    assert np.array_equal(adj.col, col)


def test_refine_jobs_alone_and_in_the_mix(gpu_ctx):
    """Refinement jobs (the clustered loop, state kept on the device across its
    rounds) submitted among graph jobs: each refinement gives Context.refine's
    codebook, each graph job its single build's."""
    rng = np.random.default_rng(31)
    ra = W.vt_refine(seed=6, total=40000)
    rb = W.vt_refine(seed=7, total=25000)
    kw = dict(kmeans_iters=5, max_rounds=4, converge_n=2, seed=3)
    want_a, _ = gpu_ctx.refine(ra, 5, k=32, host_threads=4, **kw)
    want_b, _ = gpu_ctx.refine(rb, 4, k=16, host_threads=4, **kw)
    graphs = [(W.filtered_uniform(rng, int(rng.integers(300, 5000)), int(rng.integers(11, 17))),
               int(rng.integers(3, 6))) for _ in range(30)]
    srv = igc.Server(0, pipelines=2, host_threads=4)
    try:
        ids = [srv.submit(s, ed) for s, ed in graphs[:10]]
        ia = srv.submit_refine(ra, 5, 32, **kw)
        ids += [srv.submit(s, ed) for s, ed in graphs[10:20]]
        ib = srv.submit_refine(rb, 4, 16, **kw)
        ids += [srv.submit(s, ed) for s, ed in graphs[20:]]
        res = [srv.wait(j) for j in ids]
        got_a, got_b = srv.wait(ia), srv.wait(ib)
    finally:
        srv.close()
    assert got_a["status"] == 0 and got_b["status"] == 0, (got_a["error"], got_b["error"])
    assert np.array_equal(got_a["order"], want_a) and np.array_equal(got_b["order"], want_b)
    assert got_a["batch_jobs"] == 1 and got_b["batch_jobs"] == 1
    for (s, ed), r in zip(graphs, res):
        assert r["status"] == 0
        assert np.array_equal(r["order"], igc.greedy_order(gpu_ctx.graph_build(s, ed)))
