"""The refinement loop's captured round: igc_refine replays one hipGraph per
round (gather the round's strings from the device-resident set, pack, seed,
Lloyd iterations, final assignment; grids for the first round, the round's
size read from device memory).  Every replay, with the loop's own eager work
(batched builds, copies, host greedy) between replays, must give what the same
kernels launched directly give: the same clusters, codewords and final
codebook, round by round (the 500,000-string refinement workload and a
smaller set that converges)."""
import numpy as np
import pytest

import indexgen_amd as igc
from indexgen_amd import workloads as W

pytestmark = pytest.mark.gpu


def _refine(ctx, syms, replay, **kw):
    ctx.set_graph_replay(replay)
    try:
        return ctx.refine(syms, 5, **kw)
    finally:
        ctx.set_graph_replay(True)


@pytest.mark.parametrize("total,k,rounds", [(500000, 256, 5), (30000, 16, 8)])
def test_replayed_rounds_equal_direct_launches(gpu_ctx, total, k, rounds):
    syms = W.vt_refine(seed=4, total=total)
    kw = dict(k=k, kmeans_iters=10, max_rounds=rounds, converge_n=2, seed=1, host_threads=16)
    og, sg = _refine(gpu_ctx, syms, True, **kw)
    oe, se = _refine(gpu_ctx, syms, False, **kw)
    print(f"[refine] N={total} rounds={len(sg['rounds'])} converged={sg['converged']} codewords={og.size}",
          flush=True)
    assert len(sg["rounds"]) >= 3
    key = lambda st: [(r["candidates"], r["clusters"], r["codewords"], r["pairs"]) for r in st["rounds"]]
    assert key(sg) == key(se)
    assert sg["converged"] == se["converged"]
    assert np.array_equal(og, oe)
    # a second call on the same context captures its own graph and gives the same
    og2, _ = _refine(gpu_ctx, syms, True, **kw)
    assert np.array_equal(og, og2)
