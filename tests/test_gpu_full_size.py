"""The benchmark workloads at full size against the compiled reference.

tests/golden/full_size_golden.npz holds, per set, what the reference's own
FillAdjList (src/SparseMat.cpp:567-615) and Codebook (:679-731) produce with 8
threads (tests/golden/make_golden.py full): 2|E| ("Number of Ones in Matrix"),
the SHA-256 of the sorted i<j edge list and of the codebook in selection order.
The GPU builds the same seeded sets (their symbol digests are checked first) and
must give the same three values: the edge set exactly, and the codebook under
the neighbour order of an 8-thread fill (IGC_ORDER_THREADS(8)).
Sets: the n=11 plain set (14,409 strings), the n=13 linear-code set (200,000)
and all 64 uneven sets of the batched workload (one batched build)."""
import hashlib
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import indexgen_amd as igc
from indexgen_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "full_size_golden.npz")


@pytest.fixture(scope="module")
def full():
    z = np.load(GOLDEN)
    return {k: z[k] for k in z.files}


def _check_inputs(full, name, syms, ed):
    assert tuple(full[f"{name}_shape"]) == syms.shape, name
    assert int(full[f"{name}_min_ed"]) == ed, name
    got = hashlib.sha256(np.ascontiguousarray(syms).tobytes()).hexdigest()
    assert got == full[f"{name}_syms_sha256"].item().decode(), f"{name}: workload generator drifted"


def _check_set(full, name, syms, upper, ordered, order):
    T = int(full["threads"])
    assert 2 * upper.n_edges == int(full[f"{name}_ones"]), name
    assert 2 * ordered.n_edges == int(full[f"{name}_ones"]), name
    assert igc.edges_sha256(upper) == full[f"{name}_edges_sha256"].item().decode(), f"{name}: edge set"
    assert order.size == int(full[f"{name}_codewords"]), f"{name}: codebook size (T={T})"
    assert igc.codebook_sha256(syms, order) == full[f"{name}_codebook_sha256"].item().decode(), f"{name}: codebook"


@pytest.mark.parametrize("name", ["plain", "linear"])
def test_single_sets_match_reference(gpu_ctx, full, name):
    syms, ed = (W.plain_small(seed=1), 3) if name == "plain" else (W.linear_large(seed=2), 4)
    _check_inputs(full, name, syms, ed)
    T = int(full["threads"])
    upper = gpu_ctx.graph_build(syms, ed, igc.IGC_UPPER_TRIANGLE)
    ordered = gpu_ctx.graph_build(syms, ed, igc.order_threads(T))
    print(f"[full] {name}: N={syms.shape[0]} edges={upper.n_edges}", flush=True)
    _check_set(full, name, syms, upper, ordered, igc.greedy_order(ordered))


def test_batched_sets_match_reference(gpu_ctx, full):
    sets = W.batched(seed=3)
    names = [f"batched{s}" for s in range(len(sets))]
    for nm, cs in zip(names, sets):
        _check_inputs(full, nm, cs.syms, cs.edit_dist)
    T = int(full["threads"])
    pairs = [(cs.syms, cs.edit_dist) for cs in sets]
    ordered = gpu_ctx.graph_build_batch(pairs, igc.order_threads(T))
    print(f"[full] batched: {len(sets)} sets, edges={sum(a.n_edges for a in ordered)}", flush=True)
    with ThreadPoolExecutor(8) as ex:  # the greedy runs in the library, outside the GIL
        orders = list(ex.map(igc.greedy_order, ordered))
    upper = gpu_ctx.graph_build_batch(pairs, igc.IGC_UPPER_TRIANGLE)
    for nm, cs, u, o, od in zip(names, sets, upper, ordered, orders):
        _check_set(full, nm, cs.syms, u, o, od)
