"""GPU parity of the on-device candidate generators (LinearCode, VTCode,
DifferentialVTCode, AllStrings, each with applyFilters) with the lists the
compiled reference's generators wrote (tests/golden/gen_golden.npz), with the
oracle on further parameters, and, at sizes the oracle cannot enumerate in
seconds, with counting identities and the enumeration order."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from test_gen_oracle import GOLD, cases, check_kept, digest, filt

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def host(ctx, c):
    return ctx.candidates_to_host(c)


@pytest.mark.parametrize("key", cases("linear") + cases("vt") + cases("dvt") + cases("all"))
def test_generators_match_reference(gpu_ctx, key):
    banner(f"generator fixture {key}")
    if key.startswith("linear"):
        full = gpu_ctx.gen_linear(GOLD[f"{key}_G"], GOLD[f"{key}_bias"])
        assert digest(host(gpu_ctx, full)) == GOLD[f"{key}_full_digest"].item().decode()
        c = gpu_ctx.gen_linear(GOLD[f"{key}_G"], GOLD[f"{key}_bias"], *filt(key))
    elif key.startswith("vt"):
        n, a, b = (int(x) for x in GOLD[f"{key}_params"])
        c = gpu_ctx.gen_vt(n, a, b, *filt(key))
    elif key.startswith("dvt"):
        n, s = (int(x) for x in GOLD[f"{key}_params"])
        c = gpu_ctx.gen_diff_vt(n, s, *filt(key))
    else:
        (n,) = (int(x) for x in GOLD[f"{key}_params"])
        c = gpu_ctx.gen_all(n, *filt(key))
    check_kept(key, host(gpu_ctx, c))


FILTERS = [(0, 0.0, 0.0), (1, 0.0, 0.0), (2, 0.3, 0.7), (3, 0.25, 0.75), (4, 0.5, 0.5), (0, 0.0, 0.4), (3, 0.6, 1.0)]


def test_linear_random_codes_match_oracle(gpu_ctx):
    rng = np.random.default_rng(91)
    for trial in range(24):
        k = int(rng.integers(1, 10))
        n = int(rng.integers(max(k, 2), 33))
        G = rng.integers(0, 4, size=(k, n), dtype=np.uint8)
        bias = None if trial % 3 == 0 else rng.integers(0, 4, n).astype(np.uint8)
        f = FILTERS[trial % len(FILTERS)]
        banner(f"linear k={k} n={n} filter={f}")
        got = host(gpu_ctx, gpu_ctx.gen_linear(G, bias, *f))
        want = O.gen_linear(G, bias, *f)
        assert np.array_equal(got, want), (k, n, f)


def test_enumerating_generators_match_oracle(gpu_ctx):
    rng = np.random.default_rng(92)
    for n in range(1, 11):
        for _ in range(3):
            f = FILTERS[int(rng.integers(0, len(FILTERS)))]
            a, b, s = int(rng.integers(-20, 40)), int(rng.integers(-6, 9)), int(rng.integers(-50, 90))
            banner(f"vt/dvt/all n={n} a={a} b={b} s={s} filter={f}")
            assert np.array_equal(host(gpu_ctx, gpu_ctx.gen_vt(n, a, b, *f)), O.gen_vt(n, a, b, *f)), (n, a, b, f)
            assert np.array_equal(host(gpu_ctx, gpu_ctx.gen_diff_vt(n, s, *f)), O.gen_dvt(n, s, *f)), (n, s, f)
        f = FILTERS[n % len(FILTERS)]
        assert np.array_equal(host(gpu_ctx, gpu_ctx.gen_all(n, *f)), O.gen_all(n, *f)), (n, f)


def _index(syms):
    """Base-4 index of each word (last symbol least significant)."""
    v = np.zeros(syms.shape[0], np.uint64)
    for p in range(syms.shape[1]):
        v = v * np.uint64(4) + syms[:, p].astype(np.uint64)
    return v


def test_large_enumerations_partition_and_order(gpu_ctx):
    """n = 13 (6.7e7 words per call): over a = 0..n-1 the VT codes for one b
    partition the words with symbol sum b (mod 4), 4^(n-1) of them; over
    s = 0..4n-1 the differential VT codes partition all 4^n words; every list
    is in strictly increasing index order and its words meet their syndrome."""
    n = 13
    total = 0
    for a in range(n):
        c = gpu_ctx.gen_vt(n, a, 2)
        total += c.count
        if a in (0, 7):
            w = host(gpu_ctx, c)
            idx = _index(w)
            assert (np.diff(idx.astype(np.int64)) > 0).all()
            ge = (w[:, 1:] >= w[:, :-1]) * np.arange(1, n)
            assert (ge.sum(1) % n == a).all() and (w.sum(1) % 4 == 2).all()
    assert total == 4 ** (n - 1)
    banner(f"vt partition n={n} ok ({total} words)")
    total = 0
    for s in range(4 * n):
        c = gpu_ctx.gen_diff_vt(n, s)
        total += c.count
        if s == 31:
            w = host(gpu_ctx, c).astype(np.int64)
            y = np.concatenate([(w[:, :-1] - w[:, 1:]) % 4, w[:, -1:]], axis=1)
            assert ((y * np.arange(1, n + 1)).sum(1) % (4 * n) == s).all()
            assert (np.diff(_index(w.astype(np.uint8)).astype(np.int64)) > 0).all()
    assert total == 4 ** n
    banner(f"dvt partition n={n} ok")
    # the filters on a full enumeration: counts of all strings by max run and GC
    c = gpu_ctx.gen_all(12, 3, 0.3, 0.7)
    w = host(gpu_ctx, c)
    assert c.enumerated == 4 ** 12
    keep = O.filter_indices(w, 3, 0.3, 0.7)
    assert keep.size == w.shape[0]


def test_published_codebooks_from_device_candidates(gpu_ctx, golden_dir):
    """The reference's published N11 codebooks: their candidates regenerated on
    the device from the generator matrix and bias the reference used equal the
    reference's list, and the conflict graph built from them where they lie
    has the published number of ones; the published codebook is independent
    in it."""
    for key in cases("linear")[6:]:
        c = gpu_ctx.gen_linear(GOLD[f"{key}_G"], GOLD[f"{key}_bias"], *filt(key))
        w = host(gpu_ctx, c)
        match = None
        for fn in sorted(os.listdir(golden_dir)):
            if fn.startswith("published_"):
                z = np.load(os.path.join(golden_dir, fn))
                if z["syms"].shape == w.shape and np.array_equal(z["syms"], w):
                    match = z
        assert match is not None, key
        banner(f"published {key}: N={c.count} min_ed={int(match['min_ed'])}")
        adj = gpu_ctx.graph_build_candidates(c, int(match["min_ed"]))
        assert adj.matrix_ones == int(match["ones"])
        sel = np.zeros(adj.n_nodes, bool)
        sel[match["order"]] = True
        rows = np.repeat(np.arange(adj.n_nodes), np.diff(adj.row_ptr.astype(np.int64)))
        assert not (sel[rows] & sel[adj.col]).any()


def test_generator_argument_errors(gpu_ctx):
    import indexgen_amd as igc
    with pytest.raises(igc.IGCError):
        gpu_ctx.gen_linear(np.full((3, 5), 4, np.uint8))
    with pytest.raises(igc.IGCError):
        gpu_ctx.gen_linear(np.zeros((16, 20), np.uint8))
    with pytest.raises(igc.IGCError):
        gpu_ctx.gen_vt(igc.IGC_GEN_MAX_WORD_LEN + 1, 0, 0)
    with pytest.raises(igc.IGCError):
        gpu_ctx.gen_all(0)
    c = gpu_ctx.gen_vt(5, 1, 1)
    gpu_ctx.gen_all(4)
    with pytest.raises(igc.IGCError):
        gpu_ctx.candidates_to_host(c)  # no longer the context's latest generation
