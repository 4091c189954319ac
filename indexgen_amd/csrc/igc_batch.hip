// igc_batch.hip -- many uneven candidate sets in a few launches.
//
// The sets are treated as one disjoint union of graphs on a concatenated node
// space (set s owns global ids node_off[s] .. node_off[s]+N_s-1).  One pack
// kernel packs every set (per-set length n, each set's words at its own byte
// offset: 32-bit words for n <= 16, 64-bit above).  Sets with the same pair
// predicate (Landau-Vishkin at K, Myers, every pair; on 32- or 64-bit words)
// form a group, and each group is one launch of the single-set tile code over
// the concatenated tile lists of its sets (workgroup -> set by binary search
// over per-set tile prefix sums, largest sets first so the tail is cheap),
// all into one edge list; the unchanged CSR assembly then runs once over the
// union.  Set s's CSR is the row slice node_off[s].. with ids shifted back.
// Sets large enough for the exact row filter (igc_prefilter.hpp) are sorted
// by count class first, each into its own slices of the sorted-word, count,
// permutation and window-item arrays; their tiles read the sorted words and
// emit ids through the permutation.
//
// This replaces the reference's one process + one GPU context + one launch per
// cluster (SolveIndependentSet, src/SparseMat.cpp:784-851, called from the
// cluster loop at :985-1020).
#include "igc_device.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace igc {

struct SetDesc {
    TileGeom G;          // per-set geometry (G.N = N_s)
    uint64_t tile_begin; // first tile of the set within its group's launch
    uint64_t sym_off;    // byte offset of the set's symbols in the concatenated input
    uint64_t word_off;   // byte offset of the set's packed words (8-byte aligned)
    uint32_t node_off;   // first global node id
    uint32_t n;          // string length
    int32_t K;           // Levenshtein threshold (edge iff ED <= K); kAllPairs = every pair
    uint32_t filt;       // 1: row filter (sorted words at word_off of the sorted array)
    uint32_t wbase;      // first window of the set in the item arrays
    uint32_t window;     // columns per window
};

// Device arrays of the row filter shared by all filtered sets of a batch.
struct FilterBase {
    const uint2 *feat;     // per node (node_off + sorted position)
    const uint4 *items;    // per window (wbase + w) * kItemCap
    const uint32_t *bias;
    const uint32_t *count; // per window
    const uint32_t *perm;  // per node: sorted position -> local id
};

static __device__ __forceinline__ uint32_t find_set_by(const SetDesc *__restrict__ sets, uint32_t nsets, uint64_t key,
                                                       bool by_node)
{
    uint32_t lo = 0, hi = nsets - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        const uint64_t v = by_node ? (uint64_t)sets[mid].node_off : sets[mid].tile_begin;
        if (v <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One non-empty set's symbols in the caller's device memory and their byte
// offset in the concatenated buffer (offsets ascending).
struct SymSrc {
    const uint8_t *p;
    uint64_t off;
};

// Concatenates the sets' symbol arrays in one launch (instead of one copy per
// set): thread t assembles output bytes [16t, 16t + 16) and writes them with
// one aligned 16-byte store; the tail is written byte by byte.
static __global__ void k_gather_syms(const SymSrc *__restrict__ src, uint32_t nsrc, uint64_t total,
                                     uint8_t *__restrict__ dst)
{
    const uint64_t b0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
    if (b0 >= total) return;
    uint32_t lo = 0, hi = nsrc - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (src[mid].off <= b0) lo = mid;
        else hi = mid - 1;
    }
    uint32_t s = lo;
    uint64_t next = s + 1 < nsrc ? src[s + 1].off : total;
    uint32_t w[4] = {0u, 0u, 0u, 0u};
    const uint32_t nb = (uint32_t)min<uint64_t>(16, total - b0);
    for (uint32_t q = 0; q < nb; ++q) {
        const uint64_t b = b0 + q;
        while (b >= next) {
            ++s;
            next = s + 1 < nsrc ? src[s + 1].off : total;
        }
        w[q >> 2] |= (uint32_t)src[s].p[b - src[s].off] << (8 * (q & 3));
    }
    if (nb == 16) {
        *reinterpret_cast<uint4 *>(dst + b0) = make_uint4(w[0], w[1], w[2], w[3]);
    } else {
        for (uint32_t q = 0; q < nb; ++q) dst[b0 + q] = (uint8_t)(w[q >> 2] >> (8 * (q & 3)));
    }
}

static __global__ void k_pack_batch(const uint8_t *__restrict__ syms, const SetDesc *__restrict__ sets,
                                    uint32_t nsets, uint32_t total, uint8_t *__restrict__ packed,
                                    uint32_t *__restrict__ nbase, uint32_t *__restrict__ bad)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const uint32_t s = find_set_by(sets, nsets, g, true);
    const SetDesc &d = sets[s];
    const uint32_t n = d.n;
    const uint32_t i = g - d.node_off;
    nbase[g] = d.node_off; // neighbours are stored relative to it in the CSR sort
    const uint8_t *p = syms + d.sym_off + (size_t)i * n;
    uint32_t or_all = 0;
    for (uint32_t q = 0; q < n; ++q) or_all |= p[q];
    if (n > IGC_MAX_QUATERNARY_LEN) {
        reinterpret_cast<uint64_t *>(packed + d.word_off)[i] = pack_bin64(p, n);
        if (or_all > 1u) atomicOr(bad, 2u);
        return;
    }
    if (n > 16) reinterpret_cast<uint64_t *>(packed + d.word_off)[i] = pack_word64(p, n);
    else reinterpret_cast<uint32_t *>(packed + d.word_off)[i] = pack_word(p, n);
    if (or_all > 3u) atomicOr(bad, 1u);
}

// One group: workgroup tile0 + blockIdx.x is a tile of the set whose tile
// range holds it; the set's predicate is the group's (Pol), one register of
// columns per lane.
template <class Pol>
__global__ __launch_bounds__(256) void k_pairs_sets(const uint8_t *__restrict__ packed,
                                                    const uint8_t *__restrict__ packed_s, FilterBase fb,
                                                    const SetDesc *__restrict__ sets, uint32_t nsets, uint64_t tile0,
                                                    uint64_t *__restrict__ edges, uint64_t cap,
                                                    unsigned long long *__restrict__ counter)
{
    __shared__ uint64_t stage[kWaves][kStage];
    const uint32_t lane = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t t = tile0 + blockIdx.x;
    const uint32_t s = find_set_by(sets, nsets, t, false);
    const SetDesc d = sets[s];
    uint32_t rb, cb;
    decode_tile(t - d.tile_begin, d.G, rb, cb);
    RowFilter flt{};
    if (d.filt)
        flt = filter_view(fb.feat + d.node_off, fb.items + (size_t)d.wbase * kItemCap,
                          fb.bias + (size_t)d.wbase * kItemCap, fb.count + d.wbase, fb.perm + d.node_off, d.window,
                          d.K);
    const auto *words = reinterpret_cast<const typename Pol::Word *>((d.filt ? packed_s : packed) + d.word_off);
    wave_tile<Pol, 1, true>(words, d.G, rb, cb, d.node_off, d.n, d.K, flt, stage[wave], lane, wave, edges, cap,
                            counter);
}

// Gathers row_ptr at the set boundaries so the host can split the union CSR.
static __global__ void k_gather_bounds(const uint64_t *__restrict__ row_ptr, const SetDesc *__restrict__ sets,
                                       uint32_t nsets, uint32_t total, uint64_t *__restrict__ out)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsets) out[s] = row_ptr[sets[s].node_off];
    if (s == nsets) out[s] = row_ptr[total];
}

// Sets of one pair predicate, launched together.
struct BatchGroup {
    PredKind pk;
    int K;             // Landau-Vishkin threshold (pk == LV)
    int wc;            // word class (word_class: 32-bit, 64-bit two-plane, 64-bit binary)
    uint32_t first;    // first descriptor of the group in BatchPlan::desc
    uint32_t count;    // descriptors in the group
    uint64_t tiles;    // tiles of the group's launch
};

struct BatchPlan {
    std::vector<SetDesc> desc;     // by group, largest sets first within a group; tile_begin per group
    std::vector<SetDesc> by_node;  // every set's descriptor in caller order; node_off ascending
    std::vector<BatchGroup> groups;
    std::vector<uint32_t> order;   // desc index -> caller's set index
    std::vector<uint32_t> node_off_by_set;
    std::vector<uint32_t> filtered;// desc indices with the row filter
    uint64_t windows = 0;          // filter windows of all filtered sets
    uint64_t pairs = 0;
    uint64_t sym_bytes = 0;
    uint64_t word_bytes = 0;       // packed words of all sets
    uint32_t total = 0;
    uint32_t max_set = 0;          // largest set (strings)
    uint64_t cap = 0;
};

// Plans a batch: per-set geometry (one register of columns per lane, rows per
// tile by set size), groups by predicate, filter windows, edge estimate.
static int plan_batch(igc_ctx *ctx, const igc_set *sets, uint32_t count, BatchPlan &P)
{
    P = BatchPlan();
    P.node_off_by_set.assign(count + 1, 0);
    std::vector<int> K(count);
    uint64_t total = 0, symb = 0, wordb = 0;
    std::vector<uint64_t> sym_off(count), word_off(count);
    for (uint32_t s = 0; s < count; ++s) {
        const igc_set &x = sets[s];
        if (x.n < 1 || x.n > IGC_MAX_LEN)
            return set_error(IGC_ERR_INVALID, "set %u: string length n=%u outside 1..%d", s, x.n, IGC_MAX_LEN);
        if (!x.syms && x.N) return set_error(IGC_ERR_INVALID, "set %u: null symbols", s);
        K[s] = x.edit_dist == 0 ? -1 : effective_k(x.n, x.edit_dist);
        P.node_off_by_set[s] = (uint32_t)total;
        sym_off[s] = symb;
        word_off[s] = wordb;
        total += x.N;
        P.max_set = std::max(P.max_set, x.N);
        symb += (uint64_t)x.N * x.n;
        wordb += ((uint64_t)x.N * word_bytes(x.n) + 7) & ~7ull;
        P.pairs += (uint64_t)x.N * (x.N > 0 ? x.N - 1 : 0) / 2;
    }
    if (total > 0x7FFFFFFFull) return set_error(IGC_ERR_INVALID, "batch has %llu nodes", (unsigned long long)total);
    P.node_off_by_set[count] = (uint32_t)total;
    P.total = (uint32_t)total;
    P.sym_bytes = symb;
    P.word_bytes = wordb;
    P.by_node.resize(count);
    // group key: predicate kind, K for Landau-Vishkin, word class
    auto key = [&](uint32_t s) -> int {
        const PredKind pk = ctx_pred(ctx, sets[s].n, K[s]);
        const int k = pk == PredKind::LV ? K[s] : 0;
        return ((int)pk * 64 + k) * 3 + word_class(sets[s].n);
    };
    std::vector<uint32_t> run; // runnable sets, by group then size (largest first)
    for (uint32_t s = 0; s < count; ++s) {
        SetDesc d{};
        const igc_set &x = sets[s];
        d.G = choose_geometry(x.N, lane_cols(x.n), x.N >= 8192 ? 1u : (x.N >= 2048 ? 2u : 4u));
        d.sym_off = sym_off[s];
        d.word_off = word_off[s];
        d.node_off = P.node_off_by_set[s];
        d.n = x.n;
        d.K = K[s];
        P.by_node[s] = d;
        if (K[s] >= 0 && x.N >= 2) run.push_back(s);
    }
    std::stable_sort(run.begin(), run.end(), [&](uint32_t a, uint32_t b) {
        const int ka = key(a), kb = key(b);
        if (ka != kb) return ka < kb;
        return sets[a].N > sets[b].N;
    });
    double est_edges = 0;
    for (size_t r = 0; r < run.size(); ++r) {
        const uint32_t s = run[r];
        const igc_set &x = sets[s];
        SetDesc d = P.by_node[s];
        if (r == 0 || key(s) != key(run[r - 1])) {
            const PredKind pk = ctx_pred(ctx, x.n, K[s]);
            P.groups.push_back(BatchGroup{pk, pk == PredKind::LV ? K[s] : 0, word_class(x.n), (uint32_t)P.desc.size(), 0, 0});
        }
        BatchGroup &g = P.groups.back();
        if (filter_pays(ctx->prefilter, x.N, x.n, K[s])) {
            d.filt = 1;
            d.wbase = (uint32_t)P.windows;
            d.window = 64u * (uint32_t)lane_cols(x.n);
            P.windows += (x.N + d.window - 1) / d.window;
            P.filtered.push_back((uint32_t)P.desc.size());
        }
        d.tile_begin = g.tiles;
        g.tiles += tile_count(d.G);
        g.count++;
        est_edges += (double)initial_edge_cap(x.N, x.n, K[s]);
        P.by_node[s] = d;
        P.order.push_back(s);
        P.desc.push_back(d);
    }
    P.cap = ctx->first_edge_cap ? ctx->first_edge_cap : std::max<uint64_t>((uint64_t)est_edges, 1);
    if (P.windows > 0xFFFFFFFFull) return set_error(IGC_ERR_INVALID, "batch has too many filter windows");
    return IGC_OK;
}

template <class Pol>
static void launch_sets_t(hipStream_t s, const uint8_t *packed, const uint8_t *packed_s, const FilterBase &fb,
                          const SetDesc *d_sets, uint32_t nsets, uint64_t tile0, uint32_t ntiles, uint64_t *edges,
                          uint64_t cap, unsigned long long *counter)
{
    hipLaunchKernelGGL((k_pairs_sets<Pol>), dim3(ntiles), dim3(256), 0, s, packed, packed_s, fb, d_sets, nsets, tile0,
                       edges, cap, counter);
}

// Policies of one word class: two-plane words of type W, or binary words (BIN).
template <class W, bool BIN> struct SetPols {
    template <int K> using LV = typename LvPolOf<K, W, false>::type;
    using Myers = PolMyers<W>;
    using All = PolAll<W>;
};
template <class W> struct SetPols<W, true> {
    template <int K> using LV = PolLVBin<K>;
    using Myers = PolMyersBin;
    using All = PolAllBin;
};

template <class W, bool BIN = false>
static int launch_sets_w(const BatchGroup &g, hipStream_t s, const uint8_t *packed, const uint8_t *packed_s,
                         const FilterBase &fb, const SetDesc *d_sets, uint64_t tile0, uint32_t ntiles,
                         uint64_t *edges, uint64_t cap, unsigned long long *counter)
{
    using Pols = SetPols<W, BIN>;
    if (g.pk == PredKind::MYERS) {
        launch_sets_t<typename Pols::Myers>(s, packed, packed_s, fb, d_sets, g.count, tile0, ntiles, edges, cap,
                                            counter);
        return IGC_OK;
    }
    if (g.pk == PredKind::ALL) {
        launch_sets_t<typename Pols::All>(s, packed, packed_s, fb, d_sets, g.count, tile0, ntiles, edges, cap,
                                          counter);
        return IGC_OK;
    }
    switch (g.K) {
#define IGC_GCASE(k)                                                                                               \
    case k:                                                                                                        \
        launch_sets_t<typename Pols::template LV<k>>(s, packed, packed_s, fb, d_sets, g.count, tile0, ntiles,      \
                                                     edges, cap, counter);                                         \
        return IGC_OK;
        IGC_GCASE(0) IGC_GCASE(1) IGC_GCASE(2) IGC_GCASE(3) IGC_GCASE(4) IGC_GCASE(5) IGC_GCASE(6) IGC_GCASE(7)
        IGC_GCASE(8) IGC_GCASE(9) IGC_GCASE(10) IGC_GCASE(11) IGC_GCASE(12) IGC_GCASE(13)
#undef IGC_GCASE
    default:
        return set_error(IGC_ERR_INVALID, "no Landau-Vishkin kernel for K=%d", g.K);
    }
}

// The group's tiles in launches of at most 2^31 - 1 workgroups, on stream st.
static int launch_group(igc_ctx *ctx, hipStream_t st, const BatchGroup &g, const FilterBase &fb,
                        const SetDesc *d_group, uint64_t *edges, uint64_t cap, unsigned long long *counter,
                        uint32_t *launches)
{
    constexpr uint64_t kMaxGrid = 0x7FFFFFFFull;
    for (uint64_t t0 = 0; t0 < g.tiles; t0 += kMaxGrid) {
        const uint32_t nt = (uint32_t)std::min<uint64_t>(g.tiles - t0, kMaxGrid);
        const uint8_t *pk = ctx->packed.as<uint8_t>(), *pks = ctx->packed_s.as<uint8_t>();
        const int rc =
            g.wc == 2   ? launch_sets_w<uint64_t, true>(g, st, pk, pks, fb, d_group, t0, nt, edges, cap, counter)
            : g.wc == 1 ? launch_sets_w<uint64_t>(g, st, pk, pks, fb, d_group, t0, nt, edges, cap, counter)
                        : launch_sets_w<uint32_t>(g, st, pk, pks, fb, d_group, t0, nt, edges, cap, counter);
        if (rc) return rc;
        HIP_TRY(hipGetLastError());
        ++*launches;
    }
    return IGC_OK;
}

// All groups' launches.  With several groups, group g > 0 runs on auxiliary
// stream (g - 1) mod kAuxStreams (default priority), forked from and joined
// back to the context's stream by events, so one group's tail of workgroups
// overlaps the next group instead of leaving CUs idle between launches.  All
// launches append to the same edge list through its atomic counter.
static int launch_groups(igc_ctx *ctx, const BatchPlan &P, const FilterBase &fb, const SetDesc *d_groups,
                         uint64_t cap, unsigned long long *counter, uint32_t *launches)
{
    hipStream_t s = ctx->stream;
    uint64_t *edges = ctx->edges.as<uint64_t>();
    const size_t ng = P.groups.size();
    const uint32_t naux = (uint32_t)std::min<size_t>(ng > 0 ? ng - 1 : 0, igc_ctx::kAuxStreams);
    int rc;
    for (uint32_t a = 0; a < naux; ++a) {
        if (!ctx->aux[a]) HIP_TRY(hipStreamCreateWithFlags(&ctx->aux[a], hipStreamNonBlocking));
        if (!ctx->aux_ev[a]) HIP_TRY(hipEventCreateWithFlags(&ctx->aux_ev[a], hipEventDisableTiming));
    }
    if (naux) {
        HIP_TRY(hipEventRecord(ctx->aux_ev[0], s)); // fork point: pack and filter set-up done
        for (uint32_t a = 0; a < naux; ++a) HIP_TRY(hipStreamWaitEvent(ctx->aux[a], ctx->aux_ev[0], 0));
    }
    for (size_t gi = 0; gi < ng; ++gi) {
        const BatchGroup &g = P.groups[gi];
        hipStream_t st = gi == 0 ? s : ctx->aux[(gi - 1) % igc_ctx::kAuxStreams];
        if ((rc = launch_group(ctx, st, g, fb, d_groups + g.first, edges, cap, counter, launches))) return rc;
    }
    for (uint32_t a = 0; a < naux; ++a) { // join
        HIP_TRY(hipEventRecord(ctx->aux_ev[a], ctx->aux[a]));
        HIP_TRY(hipStreamWaitEvent(s, ctx->aux_ev[a], 0));
    }
    return IGC_OK;
}

// Output buffers of the row filter for the batch's filtered sets, all sized
// before the first set-up launch (no reallocation while kernels run).
static int batch_filter_buffers(igc_ctx *ctx, const BatchPlan &P, FilterBase &fb)
{
    fb = FilterBase{};
    if (P.filtered.empty()) return IGC_OK;
    int rc;
    if ((rc = ctx->feat_s.ensure(sizeof(uint2) * (size_t)P.total))) return rc;
    if ((rc = ctx->perm.ensure(sizeof(uint32_t) * (size_t)P.total))) return rc;
    if ((rc = ctx->packed_s.ensure(std::max<size_t>(P.word_bytes, 8)))) return rc;
    if ((rc = ctx->fitems.ensure(sizeof(uint4) * kItemCap * (size_t)P.windows))) return rc;
    if ((rc = ctx->fbias.ensure(sizeof(uint32_t) * kItemCap * (size_t)P.windows))) return rc;
    if ((rc = ctx->fcount.ensure(sizeof(uint32_t) * (size_t)P.windows))) return rc;
    fb.feat = ctx->feat_s.as<uint2>();
    fb.items = ctx->fitems.as<uint4>();
    fb.bias = ctx->fbias.as<uint32_t>();
    fb.count = ctx->fcount.as<uint32_t>();
    fb.perm = ctx->perm.as<uint32_t>();
    return IGC_OK;
}

static int batch_device_locked(igc_ctx *ctx, const uint8_t *d_syms, const BatchPlan &Pin,
                               std::vector<uint64_t> &bounds_out, uint64_t *n_edges)
{
    BatchPlan P = Pin;
    hipStream_t s = ctx->stream;
    igc_timing tm{};
    tm.pairs = P.pairs;
    const uint32_t nr = (uint32_t)P.desc.size();    // runnable sets, by group
    const uint32_t nset = (uint32_t)P.by_node.size(); // every set, caller order
    int rc;
    if ((rc = ctx->packed.ensure(std::max<size_t>(P.word_bytes, 8)))) return rc;
    if ((rc = ctx->misc.ensure(sizeof(uint32_t) * kMiscWords))) return rc;
    if ((rc = ctx->sets.ensure(sizeof(SetDesc) * std::max<size_t>((size_t)nr + nset, 1)))) return rc;
    if ((rc = ctx->bounds.ensure(sizeof(uint64_t) * ((size_t)nset + 1)))) return rc;
    if ((rc = ctx->nbase.ensure(sizeof(uint32_t) * std::max<size_t>(P.total, 1)))) return rc;
    const SetDesc *d_groups = ctx->sets.as<SetDesc>();
    const SetDesc *d_node_order = ctx->sets.as<SetDesc>() + nr;
    if (nr) HIP_TRY(hipMemcpyAsync(ctx->sets.p, P.desc.data(), sizeof(SetDesc) * nr, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->sets.as<SetDesc>() + nr, P.by_node.data(), sizeof(SetDesc) * nset,
                           hipMemcpyHostToDevice, s));
    FilterBase fb;
    if ((rc = batch_filter_buffers(ctx, P, fb))) return rc;
    uint32_t *misc = ctx->misc.as<uint32_t>();
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(misc);
    uint64_t cap = P.cap;
    for (int attempt = 0; attempt < 3; ++attempt) {
        if ((rc = ctx->edges.ensure(sizeof(uint64_t) * (size_t)cap))) return rc;
        HIP_TRY(hipEventRecord(ctx->ev[0], s));
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, misc, (uint64_t)kMiscWords, 0u);
        if (P.total)
            hipLaunchKernelGGL(k_pack_batch, dim3((P.total + 255) / 256), dim3(256), 0, s, d_syms, d_node_order, nset,
                               P.total, ctx->packed.as<uint8_t>(), ctx->nbase.as<uint32_t>(), misc + 2);
        HIP_TRY(hipGetLastError());
        uint32_t launches = 2;
        if (!P.filtered.empty()) {
            std::vector<FilterSet> fs(P.filtered.size());
            for (size_t q = 0; q < P.filtered.size(); ++q) {
                const SetDesc &d = P.desc[P.filtered[q]];
                fs[q] = FilterSet{};
                fs[q].sym_off = d.sym_off;
                fs[q].word_off = d.word_off;
                fs[q].node_off = d.node_off;
                fs[q].N = d.G.N;
                fs[q].n = d.n;
                fs[q].wbase = d.wbase;
                fs[q].window = d.window;
            }
            if ((rc = filter_sets(ctx, d_syms, ctx->packed.p, ctx->packed_s.p, ctx->feat_s.as<uint2>(),
                                  ctx->perm.as<uint32_t>(), ctx->fitems.as<uint4>(), ctx->fbias.as<uint32_t>(),
                                  ctx->fcount.as<uint32_t>(), fs)))
                return rc;
            launches += 8;
        }
        HIP_TRY(hipEventRecord(ctx->ev[1], s));
        if ((rc = launch_groups(ctx, P, fb, d_groups, cap, counter, &launches))) return rc;
        HIP_TRY(hipEventRecord(ctx->ev[2], s));
        uint32_t host_misc[kMiscWords];
        HIP_TRY(hipMemcpyAsync(host_misc, misc, sizeof host_misc, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (host_misc[2]) return bad_symbols(host_misc[2]);
        const uint64_t E = (uint64_t)host_misc[0] | ((uint64_t)host_misc[1] << 32);
        if (E > cap) {
            cap = E + E / 64 + 1024; // exact regrow, then replay the pairs
            continue;
        }
        if ((rc = launch_csr(ctx, P.total, E, counter, misc, &launches, false, P.max_set, ctx->nbase.as<uint32_t>())))
            return rc;
        hipLaunchKernelGGL(k_gather_bounds, dim3((nset + 1 + 255) / 256), dim3(256), 0, s, ctx->row_ptr.as<uint64_t>(),
                           d_node_order, nset, P.total, ctx->bounds.as<uint64_t>());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ctx->ev[3], s));
        std::vector<uint64_t> b(nset + 1);
        HIP_TRY(hipMemcpyAsync(b.data(), ctx->bounds.p, sizeof(uint64_t) * (nset + 1), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipEventElapsedTime(&tm.pack_ms, ctx->ev[0], ctx->ev[1]));
        HIP_TRY(hipEventElapsedTime(&tm.pairs_ms, ctx->ev[1], ctx->ev[2]));
        HIP_TRY(hipEventElapsedTime(&tm.csr_ms, ctx->ev[2], ctx->ev[3]));
        HIP_TRY(hipEventElapsedTime(&tm.total_ms, ctx->ev[0], ctx->ev[3]));
        tm.launches = launches + 1;
        tm.retries = (uint32_t)attempt;
        tm.n_edges = E;
        bounds_out = b; // caller order: the gather walked the node-ordered table
        *n_edges = E;
        ctx->last = tm;
        return IGC_OK;
    }
    return set_error(IGC_ERR_DEVICE, "edge buffer kept overflowing");
}

// Host arrays of the set whose nodes start at off (row pointers from the union
// row_ptr already on the host in rp; the col entries are filled by
// d2h_split_cols).
static int alloc_set(const uint64_t *rp, uint32_t off, uint32_t N, igc_csr *out)
{
    const uint64_t base = rp[off];
    const uint64_t all = rp[off + N] - base;
    out->n_nodes = N;
    out->location = IGC_ON_HOST;
    out->n_edges = all / 2;
    out->row_ptr = static_cast<uint64_t *>(host_alloc(sizeof(uint64_t) * ((size_t)N + 1)));
    out->col = static_cast<uint32_t *>(host_alloc(sizeof(uint32_t) * std::max<uint64_t>(all, 1)));
    if (!out->row_ptr || !out->col) {
        igc_csr_free(out);
        return set_error(IGC_ERR_OOM, "host CSR allocation failed");
    }
    for (uint32_t x = 0; x <= N; ++x) out->row_ptr[x] = rp[off + x] - base;
    return IGC_OK;
}

// Streams the union col (total entries) to the sets' host CSRs through the
// pinned staging ring (d2h_stream: one transfer per chunk whatever the number
// of sets, chunks distributed on worker threads), no pinned buffer of the
// union's size.  Set slices are contiguous and ascending in caller order; each
// entry goes to its set less the set's first node.
static int d2h_split_cols(igc_ctx *ctx, const uint64_t *rp, const uint32_t *node_off, const igc_set *sets,
                          uint32_t count, uint64_t total, igc_csr *outs)
{
    if (total == 0) return IGC_OK;
    std::vector<uint64_t> ends(count); // end of each set's slice (ascending)
    for (uint32_t t = 0; t < count; ++t) ends[t] = rp[node_off[t] + sets[t].N];
    return d2h_stream(ctx, ctx->col.p, sizeof(uint32_t) * total,
                      [&](const uint8_t *chunk, size_t off, size_t bytes) {
                          const uint64_t c0 = off / sizeof(uint32_t), c1 = c0 + bytes / sizeof(uint32_t);
                          const uint32_t *in = reinterpret_cast<const uint32_t *>(chunk);
                          uint32_t t = (uint32_t)(std::upper_bound(ends.begin(), ends.end(), c0) - ends.begin());
                          for (; t < count; ++t) {
                              const uint64_t b = rp[node_off[t]], e = ends[t];
                              if (b >= c1) break;
                              const uint64_t lo = std::max(b, c0), hi = std::min(e, c1);
                              uint32_t *out = outs[t].col + (lo - b);
                              const uint32_t *from = in + (lo - c0);
                              const uint32_t shift = node_off[t];
                              for (uint64_t q = 0; q < hi - lo; ++q) out[q] = from[q] - shift;
                          }
                      });
}

// Applies the set's flags to its host CSR: IGC_UPPER_TRIANGLE keeps each row's
// neighbours above the row (in place: the write index never passes the read
// index); a thread order is applied as in the single-set call.
static void finish_set(uint32_t flags, igc_csr *out)
{
    const uint32_t N = out->n_nodes;
    if (flags & IGC_UPPER_TRIANGLE) {
        uint64_t w = 0;
        for (uint32_t x = 0; x < N; ++x) {
            const uint64_t q0 = out->row_ptr[x], q1 = out->row_ptr[x + 1];
            out->row_ptr[x] = w;
            for (uint64_t q = q0; q < q1; ++q)
                if (out->col[q] > x) out->col[w++] = out->col[q];
        }
        out->row_ptr[N] = w;
        return;
    }
    const uint32_t T = (flags >> 16) & 0xFFFFu;
    if (T > 1) reorder_threads(*out, T);
}

} // namespace igc

using namespace igc;

extern "C" {

IGC_API int igc_graph_build_batch(igc_ctx *ctx, const igc_set *sets, uint32_t count, igc_csr *outs)
{
    if (!ctx || (!sets && count) || (!outs && count)) return set_error(IGC_ERR_INVALID, "null argument");
    for (uint32_t s = 0; s < count; ++s) std::memset(&outs[s], 0, sizeof(igc_csr));
    if (count == 0) return IGC_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return batch_build_host_locked(ctx, sets, count, outs);
}

} // extern "C"

namespace igc {

// Copies the sets' host symbol arrays, concatenated, to dst through two pinned
// 8 MiB chunks: one DMA per chunk instead of one pageable copy per set (a
// batch of a server may hold hundreds of small jobs).  A chunk buffer is
// refilled only after its previous copy has completed.
static int h2d_concat_staged(igc_ctx *ctx, uint8_t *dst, const igc_set *sets, uint32_t count, uint64_t total)
{
    constexpr size_t kChunk = 8u << 20;
    if (total == 0) return IGC_OK;
    int rc = ctx->stage.ensure(2 * kChunk);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    uint32_t set = 0;
    uint64_t in_set = 0, done = 0;
    for (uint32_t c = 0; done < total; ++c) {
        uint8_t *buf = ctx->stage.as<uint8_t>() + (c & 1) * kChunk;
        if (c >= 2) HIP_TRY(hipEventSynchronize(ctx->ev[c & 1]));
        const uint64_t want = std::min<uint64_t>(kChunk, total - done);
        uint64_t fill = 0;
        while (fill < want && set < count) {
            const uint64_t bytes = (uint64_t)sets[set].N * sets[set].n;
            const uint64_t take = std::min<uint64_t>(bytes - in_set, want - fill);
            if (take) std::memcpy(buf + fill, sets[set].syms + in_set, take);
            fill += take;
            in_set += take;
            if (in_set == bytes) {
                ++set;
                in_set = 0;
            }
        }
        if (fill != want) return set_error(IGC_ERR_INVALID, "set sizes changed during the copy");
        HIP_TRY(hipMemcpyAsync(dst + done, buf, fill, hipMemcpyHostToDevice, st));
        HIP_TRY(hipEventRecord(ctx->ev[c & 1], st));
        done += fill;
    }
    return IGC_OK;
}

// The batch built from its sets' symbols concatenated at d_syms (device
// memory, set order), every set's CSR copied to host arrays.
static int batch_to_host_locked(igc_ctx *ctx, const uint8_t *d_syms, const BatchPlan &P, const igc_set *sets,
                                uint32_t count, igc_csr *outs)
{
    std::vector<uint64_t> bounds;
    uint64_t E = 0;
    int rc;
    if ((rc = batch_device_locked(ctx, d_syms, P, bounds, &E))) return rc;
    // union row_ptr -> pinned host memory; each set's col slice straight into its own host CSR
    if ((rc = ctx->host_rp.ensure(sizeof(uint64_t) * ((size_t)P.total + 1)))) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->host_rp.p, ctx->row_ptr.p, sizeof(uint64_t) * ((size_t)P.total + 1),
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const uint64_t *rp = ctx->host_rp.as<uint64_t>();
    for (uint32_t s = 0; s < count; ++s) {
        if ((rc = alloc_set(rp, P.node_off_by_set[s], sets[s].N, &outs[s]))) {
            for (uint32_t t = 0; t < s; ++t) igc_csr_free(&outs[t]);
            return rc;
        }
    }
    if ((rc = d2h_split_cols(ctx, rp, P.node_off_by_set.data(), sets, count, rp[P.total], outs))) {
        for (uint32_t t = 0; t < count; ++t) igc_csr_free(&outs[t]);
        return rc;
    }
    for (uint32_t s = 0; s < count; ++s) finish_set(sets[s].flags, &outs[s]);
    return IGC_OK;
}

int batch_build_host_locked(igc_ctx *ctx, const igc_set *sets, uint32_t count, igc_csr *outs)
{
    for (uint32_t s = 0; s < count; ++s) std::memset(&outs[s], 0, sizeof(igc_csr));
    if (count == 0) return IGC_OK;
    BatchPlan P;
    int rc = plan_batch(ctx, sets, count, P);
    if (rc) return rc;
    if ((rc = ctx->syms.ensure(std::max<size_t>(P.sym_bytes, 1)))) return rc;
    if ((rc = h2d_concat_staged(ctx, ctx->syms.as<uint8_t>(), sets, count, P.sym_bytes))) return rc;
    return batch_to_host_locked(ctx, ctx->syms.as<uint8_t>(), P, sets, count, outs);
}

int batch_build_dev_syms_locked(igc_ctx *ctx, const uint8_t *d_syms, const igc_set *sets, uint32_t count,
                                igc_csr *outs)
{
    for (uint32_t s = 0; s < count; ++s) std::memset(&outs[s], 0, sizeof(igc_csr));
    if (count == 0) return IGC_OK;
    BatchPlan P;
    int rc = plan_batch(ctx, sets, count, P);
    if (rc) return rc;
    return batch_to_host_locked(ctx, d_syms, P, sets, count, outs);
}

} // namespace igc

extern "C" {

IGC_API int igc_graph_build_batch_device(igc_ctx *ctx, const igc_set *sets, uint32_t count, igc_csr *union_out,
                                         uint32_t *node_off, uint64_t *set_edges)
{
    if (!ctx || !sets || !union_out || !node_off || !set_edges || count == 0)
        return set_error(IGC_ERR_INVALID, "null argument or empty batch");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    BatchPlan P;
    int rc = plan_batch(ctx, sets, count, P);
    if (rc) return rc;
    if ((rc = ctx->syms.ensure(std::max<size_t>(P.sym_bytes, 16)))) return rc;
    // the sets' symbols concatenated by one gather launch; the table stays
    // alive until batch_device_locked has synchronised the stream
    std::vector<SymSrc> src;
    {
        uint64_t off = 0;
        for (uint32_t s = 0; s < count; ++s) {
            const uint64_t bytes = (uint64_t)sets[s].N * sets[s].n;
            if (bytes) src.push_back(SymSrc{sets[s].syms, off});
            off += bytes;
        }
    }
    if (!src.empty()) {
        if ((rc = ctx->srcs.ensure(sizeof(SymSrc) * src.size()))) return rc;
        HIP_TRY(hipMemcpyAsync(ctx->srcs.p, src.data(), sizeof(SymSrc) * src.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        const uint64_t threads = (P.sym_bytes + 15) / 16;
        hipLaunchKernelGGL(k_gather_syms, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, ctx->stream,
                           ctx->srcs.as<SymSrc>(), (uint32_t)src.size(), P.sym_bytes, ctx->syms.as<uint8_t>());
        HIP_TRY(hipGetLastError());
    }
    std::vector<uint64_t> bounds;
    uint64_t E = 0;
    if ((rc = batch_device_locked(ctx, ctx->syms.as<uint8_t>(), P, bounds, &E))) return rc;
    for (uint32_t s = 0; s <= count; ++s) node_off[s] = P.node_off_by_set[s];
    for (uint32_t s = 0; s < count; ++s) set_edges[s] = (bounds[s + 1] - bounds[s]) / 2;
    union_out->n_nodes = P.total;
    union_out->location = IGC_ON_DEVICE;
    union_out->n_edges = E;
    union_out->row_ptr = ctx->row_ptr.as<uint64_t>();
    union_out->col = ctx->col.as<uint32_t>();
    return IGC_OK;
}

} // extern "C"
