// igc_runtime.hip -- device orchestration and C ABI of libigc.so (gfx950).
//
// One build = pack -> pairs -> (read of the exact edge count) -> CSR assembly by
// radix sort, all on the context's stream.  If the edge buffer overflowed (the
// count is exact even then), the buffer is regrown to the exact count and the
// pairs are replayed before any CSR work.
#include "igc_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>
#include <atomic>
#include <memory>

namespace igc {

static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char *last_error() { return g_err.c_str(); }

} // namespace igc

using namespace igc;

namespace igc {

// P(Levenshtein <= K) for uniformly random filtered 4-ary strings, measured
// offline; used only to size the first edge buffer (the build is exact either
// way).  Longer strings are sparser at the same K, so n > 16 takes the n = 16
// row as an upper bound; elsewhere (n < 10, K > 5, binary strings) there is no
// estimate (negative).
static double density_prior(uint32_t n, int K)
{
    static const double tab[7][6] = {
        // K=0      1        2         3         4        5
        {2e-6, 3.6e-5, 6.2e-4, 6.7e-3, 4.6e-2, 0.19},   // n=10
        {5e-7, 9e-6, 2.1e-4, 2.6e-3, 2.0e-2, 0.10},     // n=11
        {2e-7, 2e-6, 6.6e-5, 8.5e-4, 7.6e-3, 4.5e-2},   // n=12
        {1e-7, 5e-7, 1.7e-5, 2.8e-4, 2.8e-3, 2.0e-2},   // n=13
        {1e-7, 2e-7, 3.5e-6, 9.8e-5, 1.1e-3, 8.3e-3},   // n=14
        {1e-7, 1e-7, 1.5e-6, 2.3e-5, 3.6e-4, 3.3e-3},   // n=15
        {1e-7, 1e-7, 1.5e-6, 7.5e-6, 1.2e-4, 1.3e-3},   // n=16
    };
    if (K < 0 || K > 5 || n < 10 || n > IGC_MAX_QUATERNARY_LEN) return -1.0;
    return tab[std::min<uint32_t>(n, 16) - 10][K];
}

uint64_t initial_edge_cap(uint32_t N, uint32_t n, int K)
{
    const uint64_t pairs = (uint64_t)N * (N - 1) / 2;
    if (pairs <= (1ull << 22) || K >= (int)n) return std::max<uint64_t>(pairs, 1); // K >= n: every pair
    const double prior = density_prior(n, K);
    double est = prior >= 0 ? (double)pairs * prior * 2.0 + (double)(1u << 20)
                            : (double)std::min<uint64_t>(pairs, kUnknownDensityCap);
    if (est > (double)pairs) est = (double)pairs;
    return (uint64_t)est;
}

uint64_t first_edge_cap(const igc_ctx *ctx, uint32_t N, uint32_t n, int K)
{
    return ctx->first_edge_cap ? ctx->first_edge_cap : initial_edge_cap(N, n, K);
}

TileGeom choose_geometry(uint32_t N, int cols_per_lane, uint32_t q)
{
    TileGeom G;
    G.N = N;
    G.W = 256u * (uint32_t)cols_per_lane;
    G.q = q;
    G.R = G.W / q;
    G.NBr = (N + G.R - 1) / G.R;
    G.NBc = (N + G.W - 1) / G.W;
    return G;
}

uint64_t tile_count(const TileGeom &G)
{
    uint64_t t = 0;
    for (uint32_t g = 0; g < G.NBc; ++g) {
        const uint32_t rb0 = g * G.q;
        if (rb0 >= G.NBr) break;
        const uint32_t nrb = std::min(G.q, G.NBr - rb0);
        t += (uint64_t)nrb * (G.NBc - g);
    }
    return t;
}

// Tile shape: enough workgroups to fill 256 CUs several times over.
Shape default_shape(uint32_t N)
{
    if (N >= 65536) return {4, 1};
    if (N >= 16384) return {2, N >= 32768 ? 2u : 4u};
    return {1, N >= 4096 ? 4u : 8u};
}

static int size_bucket(uint32_t N)
{
    int b = 0;
    for (uint32_t t = 4096; t < N && b < 7; t <<= 2) ++b;
    return b;
}

// Index of the kernel variant a (length, threshold) pair runs, for the table
// of autotuned shapes: Landau-Vishkin on word class c (word_class) at
// c*(kLvLaunchMaxK+1)+K, then Myers and every-pair on each word class.
int pred_variant(PredKind pk, uint32_t n, int K)
{
    const int wc = word_class(n);
    switch (pk) {
    case PredKind::LV: return wc * (kLvLaunchMaxK + 1) + K;
    case PredKind::MYERS: return 3 * (kLvLaunchMaxK + 1) + wc;
    default: return 3 * (kLvLaunchMaxK + 1) + 3 + wc;
    }
}

Shape pick_shape(igc_ctx *ctx, uint32_t n, int K, uint32_t N)
{
    const Shape t = ctx->tuned[pred_variant(ctx_pred(ctx, n, K), n, K)][size_bucket(N)];
    return t.cpl ? t : default_shape(N);
}

// Launch shape of a single-set build: with the row filter one register of
// columns per lane, since the filter decides per window of 64 lanes' columns
// (128 on 16-bit lanes) and wider windows skip far fewer rows.
static Shape build_shape(igc_ctx *ctx, uint32_t n, int K, uint32_t N, bool filtered)
{
    Shape sh = pick_shape(ctx, n, K, N);
    if (filtered) sh.cpl = 1;
    return sh;
}

template <class Pol, int CPL, bool EMIT>
static void launch_pairs_t(hipStream_t s, uint64_t tiles, const void *packed, const TileGeom &G, uint32_t node_off,
                           uint32_t n, int K, const RowFilter &f, uint64_t *edges, uint64_t cap,
                           unsigned long long *counter)
{
    hipLaunchKernelGGL((k_pairs<Pol, CPL, EMIT>), dim3((uint32_t)tiles), dim3(256), 0, s,
                       static_cast<const typename Pol::Word *>(packed), G, node_off, n, K, f, edges, cap, counter);
}

// Landau-Vishkin policy of word type W (uint32_t, uint64_t: two planes) or of
// binary words (BIN)
template <class W, int CPL, bool EMIT, bool BIN = false>
static int launch_pairs_lv(hipStream_t s, uint64_t tiles, const void *packed, const TileGeom &G, uint32_t node_off,
                           uint32_t n, int K, const RowFilter &f, uint64_t *edges, uint64_t cap,
                           unsigned long long *counter)
{
    switch (K) {
#define IGC_CASE(k)                                                                                           \
    case k:                                                                                                   \
        launch_pairs_t<typename LvPolOf<k, W, BIN>::type, CPL, EMIT>(s, tiles, packed, G, node_off, n, K, f, edges, \
                                                                     cap, counter);                           \
        return IGC_OK;
        IGC_CASE(0) IGC_CASE(1) IGC_CASE(2) IGC_CASE(3) IGC_CASE(4) IGC_CASE(5) IGC_CASE(6) IGC_CASE(7)
        IGC_CASE(8) IGC_CASE(9) IGC_CASE(10) IGC_CASE(11) IGC_CASE(12) IGC_CASE(13)
#undef IGC_CASE
    default:
        return set_error(IGC_ERR_INVALID, "no Landau-Vishkin kernel for K=%d", K);
    }
}
static_assert(kLvLaunchMaxK == 13, "launch_pairs_lv instantiates K = 0..13");

template <int CPL, bool EMIT>
static int launch_pairs_c(PredKind pk, hipStream_t s, uint64_t tiles, const void *packed, const TileGeom &G, uint32_t node_off,
                          uint32_t n, int K, const RowFilter &f, uint64_t *edges, uint64_t cap,
                          unsigned long long *counter)
{
    const int wc = word_class(n);
    switch (pk) {
    case PredKind::LV:
        if (wc == 2)
            return launch_pairs_lv<uint64_t, CPL, EMIT, true>(s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
        if (wc == 1) return launch_pairs_lv<uint64_t, CPL, EMIT>(s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
        return launch_pairs_lv<uint32_t, CPL, EMIT>(s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
    case PredKind::MYERS:
        if (wc == 2) launch_pairs_t<PolMyersBin, CPL, EMIT>(s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
        else if (wc == 1) launch_pairs_t<PolMyers<uint64_t>, CPL, EMIT>(s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
        else launch_pairs_t<PolMyers<uint32_t>, CPL, EMIT>(s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
        return IGC_OK;
    default:
        if (wc == 2) launch_pairs_t<PolAllBin, CPL, EMIT>(s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
        else if (wc == 1) launch_pairs_t<PolAll<uint64_t>, CPL, EMIT>(s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
        else launch_pairs_t<PolAll<uint32_t>, CPL, EMIT>(s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
        return IGC_OK;
    }
}

int launch_pairs(PredKind pk, uint32_t n, int K, int cpl, bool emit, hipStream_t s, uint64_t tiles, const void *packed,
                 const TileGeom &G, uint32_t node_off, uint64_t *edges, uint64_t cap, unsigned long long *counter,
                 const RowFilter &f)
{
    if (tiles == 0) return IGC_OK;
    if (tiles > 0x7FFFFFFFull) return set_error(IGC_ERR_INVALID, "too many tiles (%llu)", (unsigned long long)tiles);
    int rc;
    if (cpl == 1) rc = emit ? launch_pairs_c<1, true>(pk, s, tiles, packed, G, node_off, n, K, f, edges, cap, counter)
                            : launch_pairs_c<1, false>(pk, s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
    else if (cpl == 2) rc = emit ? launch_pairs_c<2, true>(pk, s, tiles, packed, G, node_off, n, K, f, edges, cap, counter)
                                 : launch_pairs_c<2, false>(pk, s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
    else if (cpl == 4) rc = emit ? launch_pairs_c<4, true>(pk, s, tiles, packed, G, node_off, n, K, f, edges, cap, counter)
                                 : launch_pairs_c<4, false>(pk, s, tiles, packed, G, node_off, n, K, f, edges, cap, counter);
    else return set_error(IGC_ERR_INVALID, "unsupported columns-per-lane %d", cpl);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return IGC_OK;
}

// Packs N strings of length n into words of the width of n (32-bit for
// n <= 16, 64-bit above, binary words above 32); sets bit 0 of *bad when a
// symbol code exceeds 3, bit 1 when a binary string holds a code above 1.
int launch_pack(hipStream_t s, const uint8_t *d_syms, uint32_t N, uint32_t n, void *packed, uint32_t *bad)
{
    if (N == 0) return IGC_OK;
    if (binary_words(n))
        hipLaunchKernelGGL(k_pack_bin, dim3((N + 255) / 256), dim3(256), 0, s, d_syms, N, n,
                           static_cast<uint64_t *>(packed), bad);
    else if (wide_words(n))
        hipLaunchKernelGGL(k_pack<uint64_t>, dim3((N + 255) / 256), dim3(256), 0, s, d_syms, N, n,
                           static_cast<uint64_t *>(packed), bad);
    else
        hipLaunchKernelGGL(k_pack<uint32_t>, dim3((N + 255) / 256), dim3(256), 0, s, d_syms, N, n,
                           static_cast<uint32_t *>(packed), bad);
    HIP_TRY(hipGetLastError());
    return IGC_OK;
}

uint32_t grid_for(uint64_t work, uint32_t per_block, uint32_t cap_blocks)
{
    uint64_t b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap_blocks) b = cap_blocks;
    return (uint32_t)b;
}

// Effective K for the kernels: edges iff ED <= K; K >= n means every pair.
int effective_k(uint32_t n, uint32_t edit_dist)
{
    const int K = (int)edit_dist - 1;
    if (K >= (int)n) return kAllPairs;
    return K;
}

static int check_args(igc_ctx *ctx, uint32_t N, uint32_t n)
{
    if (!ctx) return set_error(IGC_ERR_INVALID, "null context");
    if (n < 1 || n > IGC_MAX_LEN) return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%d", n, IGC_MAX_LEN);
    if (N > 0x7FFFFFFFu) return set_error(IGC_ERR_INVALID, "N=%u too large", N);
    return IGC_OK;
}

// CSR of the edge list by LSD radix sort of the 2E directed entries (igc_csr.hpp),
// sized by N and cap so no host round trip is needed before the final count.
static uint32_t bits_for(uint64_t n) // smallest b with n <= 2^b (at least 1)
{
    uint32_t b = 1;
    while (b < 63 && ((uint64_t)1 << b) < n) ++b;
    return b;
}

void scan_counts(hipStream_t s, const uint32_t *counts, uint32_t L, uint64_t *bsum, uint64_t *offsets)
{
    const uint32_t nsb = (uint32_t)(((uint64_t)L + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(kScanThreads), 0, s, counts, L, bsum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, s, bsum, nsb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nsb), dim3(kScanThreads), 0, s, counts, L, bsum, offsets);
}

// Segmented CSR (igc_csr.hpp): the split pass over the edge list, then 32-bit
// passes over the low key bits.  rowbits + dbits <= 32 and the split is at
// most kRadixMaxBits bits (the caller checks).  seg_plan sizes the buffers;
// launch_csr_seg_into launches into buffers of at least those sizes.
int seg_plan(uint32_t N, uint64_t cap, bool upper, uint32_t dbits, uint32_t rowbits, uint32_t hb, SegPlan &P)
{
    P = SegPlan{};
    P.N = N;
    P.cap = cap;
    P.upper = upper;
    P.dbits = dbits;
    P.rowbits = rowbits;
    P.hb = hb;
    P.items_cap = (upper ? 1 : 2) * std::max<uint64_t>(cap, 1);
    P.nseg = 1u << hb;
    // split pass: 64-bit keys from the edge list, digit = the row's high bits
    P.tiles64 = (P.items_cap + kRadixTile - 1) / kRadixTile;
    P.ub = (P.items_cap + kSegTile - 1) / kSegTile + P.nseg; // tiles of all segments, at most
    P.lb = rowbits + dbits;                                    // low key bits
    P.passes = std::max<uint32_t>(1, (P.lb + kSegMaxBits - 1) / kSegMaxBits);
    P.bits = (P.lb + P.passes - 1) / P.passes; // at most kSegMaxBits
    if ((P.tiles64 << hb) > 0xFFFFFFFFull || (P.ub << P.bits) > 0xFFFFFFFFull)
        return set_error(IGC_ERR_INVALID, "edge buffer of %llu entries is too large", (unsigned long long)cap);
    const uint32_t L64 = (uint32_t)(P.tiles64 << hb);
    const uint32_t L32 = (uint32_t)(P.ub << P.bits);
    P.Lmax = std::max(L64, L32);
    P.slots = (size_t)P.ub * kSegTile; // padded 32-bit key slots
    P.bytes_keys0 = sizeof(uint32_t) * P.slots;
    P.bytes_keys1 = sizeof(uint32_t) * std::max<size_t>(P.slots, P.items_cap);
    P.bytes_rcount = sizeof(uint32_t) * (size_t)P.Lmax;
    P.bytes_roffset = sizeof(uint64_t) * ((size_t)P.Lmax + 1);
    P.bytes_bsum = sizeof(uint64_t) * (((size_t)P.Lmax + kScanTile - 1) / kScanTile);
    P.bytes_segmeta = sizeof(SegMeta) * P.nseg;
    return IGC_OK;
}

void launch_csr_seg_into(hipStream_t s, const SegPlan &P, const SegBufs &B, const uint64_t *edges,
                         unsigned long long *counter, uint32_t *misc, uint32_t *launches, const uint32_t *base)
{
    uint32_t *total = misc + 4; // tiles of all segments, on the device
    RadixPass rp{};
    rp.edges = edges;
    rp.cnt = counter;
    rp.cap = P.cap;
    rp.bsh = P.rowbits + P.hb;
    rp.dbits = P.dbits;
    rp.base = base;
    rp.shift = P.lb;
    rp.bits = P.hb;
    rp.from_edges = 1;
    rp.upper = P.upper ? 1u : 0u;
    const uint32_t t64 = (uint32_t)P.tiles64;
    const uint32_t L64 = (uint32_t)(P.tiles64 << P.hb);
    hipLaunchKernelGGL(k_radix_count, dim3(t64), dim3(kRadixThreads), 0, s, rp, t64, B.rcount);
    scan_counts(s, B.rcount, L64, B.bsum, B.roffset);
    hipLaunchKernelGGL(k_seg_meta, dim3(1), dim3(512), 0, s, B.roffset, t64, P.nseg, B.segmeta, total);
    if (P.hb <= 8)
        hipLaunchKernelGGL(k_radix_scatter<8>, dim3(t64), dim3(kRadixThreads), 0, s, rp, t64, B.roffset, 0u, nullptr,
                           nullptr, nullptr, B.segmeta, B.keys0);
    else
        hipLaunchKernelGGL(k_radix_scatter<kRadixMaxBits>, dim3(t64), dim3(kRadixThreads), 0, s, rp, t64, B.roffset,
                           0u, nullptr, nullptr, nullptr, B.segmeta, B.keys0);
    *launches += 6;
    // 32-bit passes, least significant digit first; the last writes col and src
    for (uint32_t pass = 0; pass < P.passes; ++pass) {
        const uint32_t shift = pass * P.bits;
        const uint32_t pb = std::min(P.bits, P.lb - shift);
        const uint32_t L = (uint32_t)(P.ub << pb);
        const bool last = pass + 1 == P.passes;
        uint32_t *in = (pass & 1) ? B.keys1 : B.keys0;
        uint32_t *out = (pass & 1) ? B.keys0 : B.keys1; // keys, or src on the last pass
        hipLaunchKernelGGL(k_seg_count, dim3((uint32_t)P.ub), dim3(kSegThreads), 0, s, in, B.segmeta, P.nseg, total,
                           shift, pb, B.rcount);
        scan_counts(s, B.rcount, L, B.bsum, B.roffset);
        hipLaunchKernelGGL(k_seg_scatter, dim3((uint32_t)P.ub), dim3(kSegThreads), 0, s, in, B.segmeta, P.nseg,
                           total, shift, pb, B.roffset, (uint32_t)last, out, P.dbits, P.rowbits, base, B.col, out);
        *launches += 5;
        if (last)
            hipLaunchKernelGGL(k_csr_row_ptr, dim3(grid_for((uint64_t)P.N + 1, 256, 8192)), dim3(256), 0, s, out,
                               counter, P.cap, P.upper ? 1u : 2u, P.N, B.row_ptr);
    }
    *launches += 1;
}

static int launch_csr_seg(igc_ctx *ctx, uint32_t N, uint64_t cap, unsigned long long *counter, uint32_t *misc,
                          uint32_t *launches, bool upper, uint32_t dbits, uint32_t rowbits, uint32_t hb,
                          const uint32_t *base)
{
    SegPlan P;
    int rc;
    if ((rc = seg_plan(N, cap, upper, dbits, rowbits, hb, P))) return rc;
    if ((rc = ctx->row_ptr.ensure(sizeof(uint64_t) * ((size_t)N + 1)))) return rc;
    if ((rc = ctx->col.ensure(sizeof(uint32_t) * P.items_cap))) return rc;
    if ((rc = ctx->keys[0].ensure(P.bytes_keys0))) return rc;
    if ((rc = ctx->keys[1].ensure(P.bytes_keys1))) return rc;
    if ((rc = ctx->rcount.ensure(P.bytes_rcount))) return rc;
    if ((rc = ctx->roffset.ensure(P.bytes_roffset))) return rc;
    if ((rc = ctx->bsum.ensure(P.bytes_bsum))) return rc;
    if ((rc = ctx->segmeta.ensure(P.bytes_segmeta))) return rc;
    SegBufs B{ctx->row_ptr.as<uint64_t>(), ctx->col.as<uint32_t>(), ctx->keys[0].as<uint32_t>(),
              ctx->keys[1].as<uint32_t>(), ctx->rcount.as<uint32_t>(), ctx->roffset.as<uint64_t>(),
              ctx->bsum.as<uint64_t>(), ctx->segmeta.as<SegMeta>()};
    launch_csr_seg_into(ctx->stream, P, B, ctx->edges.as<uint64_t>(), counter, misc, launches, base);
    HIP_TRY(hipGetLastError());
    return IGC_OK;
}

// CSR of the edge list by LSD radix sort of the 2E directed entries (igc_csr.hpp),
// sized by N and cap so no host round trip is needed before the final count.
// maxset: largest set of a batch (neighbours are stored relative to base[row],
// the first node of the row's set; 0 and null for a single set).  Large graphs
// take the segmented 32-bit form when the row's high part fits one split pass.
int launch_csr(igc_ctx *ctx, uint32_t N, uint64_t cap, unsigned long long *counter, uint32_t *misc,
               uint32_t *launches, bool upper, uint32_t maxset, const uint32_t *base)
{
    hipStream_t s = ctx->stream;
    int rc;
    const uint32_t bsh = bits_for(N); // node ids < 2^bsh
    const uint32_t dbits = base ? bits_for(std::max<uint32_t>(maxset, 2)) : bsh;
    const uint32_t rowbits = std::min(bsh, 32u - std::min(dbits, 31u));
    const uint32_t hb = bsh - rowbits;
    if (ctx->csr_form == IGC_CSR_AUTO && dbits < 32 && hb <= kRadixMaxBits)
        return launch_csr_seg(ctx, N, cap, counter, misc, launches, upper, dbits, rowbits, hb, base);
    // fewest passes of at most kRadixMaxBits-bit digits, then the narrowest digit
    const uint32_t passes = (2 * bsh + kRadixMaxBits - 1) / kRadixMaxBits;
    const uint32_t bits = (2 * bsh + passes - 1) / passes;
    const uint64_t items_cap = (upper ? 1 : 2) * std::max<uint64_t>(cap, 1);
    const uint64_t tiles64 = (items_cap + kRadixTile - 1) / kRadixTile;
    if (tiles64 << bits > 0xFFFFFFFFull)
        return set_error(IGC_ERR_INVALID, "edge buffer of %llu entries is too large", (unsigned long long)cap);
    const uint32_t tiles = (uint32_t)tiles64;
    const uint32_t L = tiles << bits; // digit-major counts
    const uint32_t nsb = (uint32_t)(((uint64_t)L + kScanTile - 1) / kScanTile);
    if ((rc = ctx->row_ptr.ensure(sizeof(uint64_t) * ((size_t)N + 1)))) return rc;
    if ((rc = ctx->col.ensure(sizeof(uint32_t) * items_cap))) return rc;
    if ((rc = ctx->keys[0].ensure(sizeof(uint64_t) * items_cap))) return rc;
    if (passes > 1 && (rc = ctx->keys[1].ensure(sizeof(uint64_t) * items_cap))) return rc;
    if ((rc = ctx->rcount.ensure(sizeof(uint32_t) * (size_t)L))) return rc;
    if ((rc = ctx->roffset.ensure(sizeof(uint64_t) * ((size_t)L + 1)))) return rc;
    if ((rc = ctx->bsum.ensure(sizeof(uint64_t) * (size_t)nsb))) return rc;
    uint32_t *counts = ctx->rcount.as<uint32_t>();
    uint64_t *offsets = ctx->roffset.as<uint64_t>();
    for (uint32_t pass = 0; pass < passes; ++pass) {
        RadixPass rp{};
        rp.keys = pass ? ctx->keys[(pass - 1) & 1].as<uint64_t>() : nullptr;
        rp.edges = ctx->edges.as<uint64_t>();
        rp.cnt = counter;
        rp.cap = cap;
        rp.bsh = bsh;
        rp.dbits = bsh;
        rp.base = nullptr;
        rp.shift = pass * bits;
        rp.bits = std::min(bits, 2 * bsh - rp.shift);
        rp.from_edges = pass == 0;
        rp.upper = upper ? 1u : 0u;
        const bool last = pass + 1 == passes;
        hipLaunchKernelGGL(k_radix_count, dim3(tiles), dim3(kRadixThreads), 0, s, rp, tiles, counts);
        hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(kScanThreads), 0, s, counts, L, ctx->bsum.as<uint64_t>());
        hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, s, ctx->bsum.as<uint64_t>(), nsb);
        hipLaunchKernelGGL(k_scan_apply, dim3(nsb), dim3(kScanThreads), 0, s, counts, L, ctx->bsum.as<uint64_t>(),
                           offsets);
        uint64_t *kout = last ? nullptr : ctx->keys[pass & 1].as<uint64_t>();
        uint32_t *ksrc = ctx->keys[pass & 1].as<uint32_t>(); // src, last pass
        if (rp.bits <= 8)
            hipLaunchKernelGGL(k_radix_scatter<8>, dim3(tiles), dim3(kRadixThreads), 0, s, rp, tiles, offsets,
                               (uint32_t)last, kout, ctx->col.as<uint32_t>(), ksrc, nullptr, nullptr);
        else
            hipLaunchKernelGGL(k_radix_scatter<kRadixMaxBits>, dim3(tiles), dim3(kRadixThreads), 0, s, rp, tiles,
                               offsets, (uint32_t)last, kout, ctx->col.as<uint32_t>(), ksrc, nullptr, nullptr);
        *launches += 5;
    }
    hipLaunchKernelGGL(k_csr_row_ptr, dim3(grid_for((uint64_t)N + 1, 256, 8192)), dim3(256), 0, s,
                       ctx->keys[(passes - 1) & 1].as<uint32_t>(), counter, cap, upper ? 1u : 2u, N,
                       ctx->row_ptr.as<uint64_t>());
    HIP_TRY(hipGetLastError());
    *launches += 1;
    return IGC_OK;
}

// Class-order table of string length n on the device (built on first use).
static int class_order_table(igc_ctx *ctx, uint32_t n, const uint32_t **out)
{
    DevBuf &t = ctx->corder[n];
    if (!t.p) {
        std::vector<uint32_t> order;
        class_order(n, order);
        int rc;
        if ((rc = t.ensure(sizeof(uint32_t) * order.size()))) return rc;
        HIP_TRY(hipMemcpyAsync(t.p, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream)); // `order` is pageable and about to go
    }
    *out = t.as<uint32_t>();
    return IGC_OK;
}

// Row-filter set-up of several sets in one set of launches (igc_prefilter.hpp):
// each set's candidates (symbols at d_syms + sym_off, words at packed +
// word_off) are sorted by count class; the sorted words go to packed_s +
// word_off, the sorted counts to feat_s[node_off..], sorted position -> local id
// to perm[node_off..], and the items of the set's windows (wbase.., `window`
// columns each) to items, bias and count.  fbase and kbase are filled in here.
// Asynchronous on ctx->stream; the table stays in ctx until the next set-up.
int filter_sets(igc_ctx *ctx, const uint8_t *d_syms, const void *packed, void *packed_s, uint2 *feat_s,
                uint32_t *perm, uint4 *items, uint32_t *bias, uint32_t *count, std::vector<FilterSet> &fs)
{
    hipStream_t s = ctx->stream;
    const uint32_t nf = (uint32_t)fs.size();
    if (nf == 0) return IGC_OK;
    uint64_t F = 0, B = 0, W = 0;
    int rc;
    for (FilterSet &f : fs) {
        const uint32_t h = f.n / 2;
        if ((rc = class_order_table(ctx, f.n, &f.order))) return rc;
        f.fbase = (uint32_t)F;
        f.kbase = (uint32_t)B;
        F += f.N;
        B += (uint64_t)comp_count(h) * comp_count(f.n - h);
        W = std::max<uint64_t>(W, (uint64_t)f.wbase + (f.N + f.window - 1) / f.window);
    }
    if (F > 0xFFFFFFFFull || B > 0xFFFFFFFFull) return set_error(IGC_ERR_INVALID, "filter set-up too large");
    const uint32_t nsb = (uint32_t)((B + kScanTile - 1) / kScanTile);
    if ((rc = ctx->feat.ensure(sizeof(uint2) * (size_t)F))) return rc;
    if ((rc = ctx->nkey.ensure(sizeof(uint32_t) * (size_t)F))) return rc;
    if ((rc = ctx->rcount.ensure(sizeof(uint32_t) * (size_t)B))) return rc;
    if ((rc = ctx->roffset.ensure(sizeof(uint64_t) * ((size_t)B + 1)))) return rc;
    if ((rc = ctx->bsum.ensure(sizeof(uint64_t) * (size_t)nsb))) return rc;
    if ((rc = ctx->fsets.ensure(sizeof(FilterSet) * nf))) return rc;
    ctx->fs_host = fs; // the copy below reads it
    HIP_TRY(hipMemcpyAsync(ctx->fsets.p, ctx->fs_host.data(), sizeof(FilterSet) * nf, hipMemcpyHostToDevice, s));
    const FilterSet *d_fs = ctx->fsets.as<FilterSet>();
    uint32_t *hist = ctx->rcount.as<uint32_t>();
    uint64_t *off = ctx->roffset.as<uint64_t>();
    const uint32_t g = (uint32_t)((F + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(B, 256, 4096)), dim3(256), 0, s, hist, B, 0u);
    hipLaunchKernelGGL(k_fs_keys, dim3(g), dim3(256), 0, s, d_syms, d_fs, nf, (uint32_t)F, ctx->feat.as<uint2>(),
                       ctx->nkey.as<uint32_t>(), hist);
    scan_counts(s, hist, (uint32_t)B, ctx->bsum.as<uint64_t>(), off);
    hipLaunchKernelGGL(k_fs_scatter, dim3(g), dim3(256), 0, s, ctx->nkey.as<uint32_t>(), d_fs, nf, (uint32_t)F, off,
                       hist, perm);
    hipLaunchKernelGGL(k_fs_permute, dim3(g), dim3(256), 0, s, d_fs, nf, (uint32_t)F, perm,
                       static_cast<const uint8_t *>(packed), ctx->feat.as<uint2>(), static_cast<uint8_t *>(packed_s),
                       feat_s);
    hipLaunchKernelGGL(k_fs_items, dim3((uint32_t)W), dim3(64), 0, s, d_fs, nf, feat_s, items, bias, count);
    HIP_TRY(hipGetLastError());
    return IGC_OK;
}

// Row filter of a single set: buffers, set-up, and the kernel's view in f;
// `window` is a wave's column span.
static int setup_filter(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n, int K, uint32_t window,
                        RowFilter &f)
{
    const uint32_t nwin = (N + window - 1) / window;
    int rc;
    if ((rc = ctx->feat_s.ensure(sizeof(uint2) * (size_t)N))) return rc;
    if ((rc = ctx->perm.ensure(sizeof(uint32_t) * (size_t)N))) return rc;
    if ((rc = ctx->packed_s.ensure(word_bytes(n) * (size_t)N))) return rc;
    if ((rc = ctx->fitems.ensure(sizeof(uint4) * kItemCap * (size_t)nwin))) return rc;
    if ((rc = ctx->fbias.ensure(sizeof(uint32_t) * kItemCap * (size_t)nwin))) return rc;
    if ((rc = ctx->fcount.ensure(sizeof(uint32_t) * (size_t)nwin))) return rc;
    std::vector<FilterSet> fs(1);
    fs[0] = FilterSet{};
    fs[0].N = N;
    fs[0].n = n;
    fs[0].window = window;
    if ((rc = filter_sets(ctx, d_syms, ctx->packed.p, ctx->packed_s.p, ctx->feat_s.as<uint2>(),
                          ctx->perm.as<uint32_t>(), ctx->fitems.as<uint4>(), ctx->fbias.as<uint32_t>(),
                          ctx->fcount.as<uint32_t>(), fs)))
        return rc;
    f = filter_view(ctx->feat_s.as<uint2>(), ctx->fitems.as<uint4>(), ctx->fbias.as<uint32_t>(),
                    ctx->fcount.as<uint32_t>(), ctx->perm.as<uint32_t>(), window, K);
    return IGC_OK;
}

// Full device build on ctx->stream.  Leaves the CSR in ctx->row_ptr/col.
static int build_device_locked(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                               uint64_t *n_edges_out, bool count_only, bool upper = false)
{
    hipStream_t s = ctx->stream;
    igc_timing tm{};
    tm.pairs = (uint64_t)N * (N > 0 ? N - 1 : 0) / 2;
    *n_edges_out = 0;
    if (edit_dist == 0 || N < 2) {
        // no pair can conflict; still produce a valid empty CSR
        int rc = ctx->row_ptr.ensure(sizeof(uint64_t) * ((size_t)N + 1));
        if (rc) return rc;
        if (!count_only) {
            hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2 * ((uint64_t)N + 1), 256, 4096)), dim3(256), 0, s,
                               ctx->row_ptr.as<uint32_t>(), 2 * ((uint64_t)N + 1), 0u);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(s));
        }
        ctx->last = tm;
        return IGC_OK;
    }
    const int K = effective_k(n, edit_dist);
    // exact row skipping by half counts: worth its sort for large sets with a
    // nontrivial threshold
    const bool use_filter = filter_pays(ctx->prefilter, N, n, K);
    const Shape sh = build_shape(ctx, n, K, N, use_filter);
    const int cpl = sh.cpl;
    const TileGeom G = choose_geometry(N, sh.cpl * lane_cols(n), sh.q);
    const uint64_t tiles = tile_count(G);

    int rc;
    if ((rc = ctx->packed.ensure(word_bytes(n) * (size_t)N))) return rc;
    if ((rc = ctx->misc.ensure(sizeof(uint32_t) * kMiscWords))) return rc;
    uint32_t *misc = ctx->misc.as<uint32_t>();
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(misc);

    uint64_t cap = count_only ? 0 : first_edge_cap(ctx, N, n, K);
    for (int attempt = 0; attempt < 3; ++attempt) {
        if (!count_only) {
            if ((rc = ctx->edges.ensure(sizeof(uint64_t) * (size_t)cap))) return rc;
        }
        HIP_TRY(hipEventRecord(ctx->ev[0], s));
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, misc, (uint64_t)kMiscWords, 0u);
        if ((rc = launch_pack(s, d_syms, N, n, ctx->packed.p, misc + 2))) return rc;
        RowFilter flt{};
        if (use_filter && (rc = setup_filter(ctx, d_syms, N, n, K, 64u * cpl * lane_cols(n), flt))) return rc;
        HIP_TRY(hipEventRecord(ctx->ev[1], s));
        rc = launch_pairs(ctx_pred(ctx, n, K), n, K, cpl, !count_only, s, tiles,
                          use_filter ? ctx->packed_s.p : ctx->packed.p, G, 0u,
                          count_only ? nullptr : ctx->edges.as<uint64_t>(), cap, counter, flt);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(ctx->ev[2], s));
        // exact edge count: sizes the CSR sort, and an overflowed attempt is
        // replayed without assembling anything
        uint32_t host_misc[kMiscWords];
        HIP_TRY(hipMemcpyAsync(host_misc, misc, sizeof host_misc, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (host_misc[2]) return bad_symbols(host_misc[2]);
        const uint64_t E = (uint64_t)host_misc[0] | ((uint64_t)host_misc[1] << 32);
        if (!count_only && E > cap) {
            cap = E + E / 64 + 1024; // exact regrow, then replay
            continue;
        }
        uint32_t launches = 3;
        if (!count_only) {
            if ((rc = launch_csr(ctx, N, E, counter, misc, &launches, upper))) return rc;
        }
        HIP_TRY(hipEventRecord(ctx->ev[3], s));
        HIP_TRY(hipEventSynchronize(ctx->ev[3]));
        HIP_TRY(hipEventElapsedTime(&tm.pack_ms, ctx->ev[0], ctx->ev[1]));
        HIP_TRY(hipEventElapsedTime(&tm.pairs_ms, ctx->ev[1], ctx->ev[2]));
        HIP_TRY(hipEventElapsedTime(&tm.csr_ms, ctx->ev[2], ctx->ev[3]));
        HIP_TRY(hipEventElapsedTime(&tm.total_ms, ctx->ev[0], ctx->ev[3]));
        tm.launches = launches;
        tm.retries = (uint32_t)attempt;
        tm.n_edges = E;
        *n_edges_out = E;
        ctx->last = tm;
        return IGC_OK;
    }
    return set_error(IGC_ERR_DEVICE, "edge buffer kept overflowing");
}

// Streams bytes of device memory at src to the host through a ring of pinned
// staging chunks.  The calling thread enqueues the chunk transfers on the
// context's stream (a ring slot is refilled once its previous chunk has been
// consumed); T worker threads take chunks round-robin, wait for their
// transfer and hand each to consume(chunk, byte offset, size) -- the host-side
// work (moving into pageable memory, which also faults its pages in, or
// splitting into per-set arrays) runs on T cores while the copy engine fills
// the next slots.  Every chunk is consumed exactly once; chunk boundaries are
// multiples of kD2HChunk.
int d2h_stream(igc_ctx *ctx, const void *src, size_t bytes, const D2HConsumer &consume)
{
    if (bytes == 0) return IGC_OK;
    const size_t chunks = (bytes + kD2HChunk - 1) / kD2HChunk;
    const uint32_t T = (uint32_t)std::min<size_t>(chunks, kD2HThreads);
    const uint32_t R = 2 * T;
    int rc = ctx->stage.ensure((size_t)R * kD2HChunk);
    if (rc) return rc;
    while (ctx->cev.size() < R) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->cev.push_back(e);
    }
    const int dev = ctx->device;
    hipStream_t st = ctx->stream;
    uint8_t *ring = ctx->stage.as<uint8_t>();
    auto size_of = [&](size_t c) { return std::min(kD2HChunk, bytes - c * kD2HChunk); };
    std::unique_ptr<std::atomic<uint8_t>[]> done(new std::atomic<uint8_t>[chunks]);
    for (size_t c = 0; c < chunks; ++c) done[c].store(0, std::memory_order_relaxed);
    std::atomic<size_t> enqueued{0};
    std::atomic<int> failed{IGC_OK};
    if (T == 1) { // one chunk: no worker thread
        HIP_TRY(hipMemcpyAsync(ring, src, bytes, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        consume(ring, 0, bytes);
        return IGC_OK;
    }
    std::vector<std::thread> workers;
    workers.reserve(T);
    for (uint32_t w = 0; w < T && failed.load() == IGC_OK; ++w) {
        try {
            workers.emplace_back([&, w] {
                (void)hipSetDevice(dev);
                for (size_t c = w; c < chunks; c += T) {
                    while (enqueued.load(std::memory_order_acquire) <= c) {
                        if (failed.load(std::memory_order_relaxed) != IGC_OK) return;
                        std::this_thread::yield();
                    }
                    if (hipEventSynchronize(ctx->cev[c % R]) != hipSuccess) {
                        failed.store(IGC_ERR_DEVICE);
                        return;
                    }
                    consume(ring + (c % R) * kD2HChunk, c * kD2HChunk, size_of(c));
                    done[c].store(1, std::memory_order_release);
                }
            });
        } catch (const std::system_error &) { // no thread: the started workers see the flag and stop
            failed.store(set_error(IGC_ERR_OOM, "staged device-to-host copy: no worker thread"));
        }
    }
    for (size_t c = 0; c < chunks && failed.load() == IGC_OK; ++c) {
        if (c >= R) // slot c % R held chunk c - R, a chunk of the same worker
            while (!done[c - R].load(std::memory_order_acquire) && failed.load() == IGC_OK) std::this_thread::yield();
        if (failed.load() != IGC_OK) break;
        const hipError_t e1 = hipMemcpyAsync(ring + (c % R) * kD2HChunk, static_cast<const uint8_t *>(src) + c * kD2HChunk,
                                             size_of(c), hipMemcpyDeviceToHost, st);
        const hipError_t e2 = e1 == hipSuccess ? hipEventRecord(ctx->cev[c % R], st) : e1;
        if (e2 != hipSuccess) {
            failed.store(set_error(IGC_ERR_DEVICE, "staged device-to-host copy failed: %s", hipGetErrorString(e2)));
            break;
        }
        enqueued.store(c + 1, std::memory_order_release);
    }
    for (auto &t : workers) t.join();
    if (failed.load() != IGC_OK) {
        (void)hipStreamSynchronize(st);
        return failed.load() == IGC_ERR_OOM ? IGC_ERR_OOM
                                            : set_error(IGC_ERR_DEVICE, "staged device-to-host copy failed");
    }
    return IGC_OK;
}

void ctx_attach_pool(igc_ctx *ctx, MemPool *pool)
{
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->mem.pool = pool;
    uint64_t held = 0;
    ctx->for_each_devbuf([&](DevBuf &b) {
        b.slot = &ctx->mem;
        held += b.p ? b.bytes : 0;
    });
    ctx->mem.grow(held); // buffers made before the attachment (none on a fresh context)
    kmeans_release(ctx);
}

void ctx_trim_locked(igc_ctx *ctx)
{
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->for_each_devbuf([](DevBuf &b) { b.release(); });
    kmeans_release(ctx);
    if (ctx->mem.pool) ctx->mem.settle();
}

} // namespace igc

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

IGC_API const char *igc_last_error(void) { return igc::last_error(); }

IGC_API const char *igc_version(void) { return "igc 0.1 gfx950"; }

IGC_API const char *igc_backend(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return "none";
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return "none";
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return "none";
    return "hip:gfx950";
}

IGC_API int igc_ctx_create(int device, igc_ctx **out)
{
    if (!out) return set_error(IGC_ERR_INVALID, "null out");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return set_error(IGC_ERR_DEVICE, "no HIP device available (libigc has no CPU fallback)");
    if (device < 0 || device >= count) return set_error(IGC_ERR_INVALID, "device %d out of range", device);
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(IGC_ERR_DEVICE, "device %d is %s, kernels are built for gfx950", device, prop.gcnArchName);
    HIP_TRY(hipSetDevice(device));
    igc_ctx *c = new (std::nothrow) igc_ctx();
    if (!c) return set_error(IGC_ERR_OOM, "host allocation failed");
    c->device = device;
    c->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return set_error(IGC_ERR_DEVICE, "hipStreamCreate failed");
    }
    c->stream = c->own_stream;
    // the no-allocation build's scratch address on this device (a failure
    // here is reported by that entry point, which retries the lookup)
    (void)scratch_attach(device);
    for (auto &e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return set_error(IGC_ERR_DEVICE, "hipEventCreate failed");
        }
    }
    *out = c;
    return IGC_OK;
}

IGC_API int igc_ctx_destroy(igc_ctx *ctx)
{
    if (!ctx) return IGC_OK;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->for_each_devbuf([](DevBuf &b) { b.release(); });
        for (HostBuf *b : {&ctx->host_rp, &ctx->stage}) b->release();
        kmeans_release(ctx);
        for (auto &e : ctx->ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : ctx->cev) (void)hipEventDestroy(e);
        for (uint32_t a = 0; a < igc_ctx::kAuxStreams; ++a) {
            if (ctx->aux[a]) (void)hipStreamSynchronize(ctx->aux[a]);
            if (ctx->aux_ev[a]) (void)hipEventDestroy(ctx->aux_ev[a]);
            if (ctx->aux[a]) (void)hipStreamDestroy(ctx->aux[a]);
        }
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
        if (ctx->mem.pool) ctx->mem.settle();
    }
    delete ctx;
    return IGC_OK;
}

IGC_API int igc_ctx_set_stream(igc_ctx *ctx, void *hip_stream)
{
    if (!ctx) return set_error(IGC_ERR_INVALID, "null context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own_stream;
    return IGC_OK;
}

IGC_API void *igc_ctx_stream(igc_ctx *ctx) { return ctx ? static_cast<void *>(ctx->stream) : nullptr; }

IGC_API int igc_ctx_set_prefilter(igc_ctx *ctx, int mode)
{
    if (!ctx) return set_error(IGC_ERR_INVALID, "null context");
    if (mode < IGC_PREFILTER_OFF || mode > IGC_PREFILTER_ALWAYS)
        return set_error(IGC_ERR_INVALID, "prefilter mode %d is not 0, 1 or 2", mode);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->prefilter = mode;
    return IGC_OK;
}

IGC_API int igc_ctx_set_predicate(igc_ctx *ctx, int predicate)
{
    if (!ctx) return set_error(IGC_ERR_INVALID, "null context");
    if (predicate != IGC_PRED_AUTO && predicate != IGC_PRED_LANDAU_VISHKIN && predicate != IGC_PRED_MYERS)
        return set_error(IGC_ERR_INVALID, "predicate %d is not IGC_PRED_AUTO, IGC_PRED_LANDAU_VISHKIN or IGC_PRED_MYERS",
                         predicate);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->predicate = predicate;
    return IGC_OK;
}

IGC_API int igc_ctx_set_csr_form(igc_ctx *ctx, int form)
{
    if (!ctx) return set_error(IGC_ERR_INVALID, "null context");
    if (form != IGC_CSR_AUTO && form != IGC_CSR_WIDE_KEYS)
        return set_error(IGC_ERR_INVALID, "CSR form %d is not 0 or 1", form);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->csr_form = form;
    return IGC_OK;
}

IGC_API int igc_ctx_set_edge_cap(igc_ctx *ctx, uint64_t edges)
{
    if (!ctx) return set_error(IGC_ERR_INVALID, "null context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->first_edge_cap = edges;
    return IGC_OK;
}

IGC_API int igc_ctx_set_graph_replay(igc_ctx *ctx, int on)
{
    if (!ctx) return set_error(IGC_ERR_INVALID, "null context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->graph_replay = on != 0;
    return IGC_OK;
}

IGC_API int igc_last_timing(igc_ctx *ctx, igc_timing *out)
{
    if (!ctx || !out) return set_error(IGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    *out = ctx->last;
    return IGC_OK;
}

IGC_API int igc_graph_build_device(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                                   igc_csr *out)
{
    int rc = check_args(ctx, N, n);
    if (rc) return rc;
    if (!out || (!d_syms && N)) return set_error(IGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    uint64_t E = 0;
    rc = build_device_locked(ctx, d_syms, N, n, edit_dist, &E, false);
    if (rc) return rc;
    out->n_nodes = N;
    out->location = IGC_ON_DEVICE;
    out->n_edges = E;
    out->row_ptr = ctx->row_ptr.as<uint64_t>();
    out->col = ctx->col.as<uint32_t>();
    return IGC_OK;
}

IGC_API int igc_csr_free(igc_csr *csr)
{
    if (!csr) return IGC_OK;
    if (csr->location == IGC_ON_HOST) {
        host_free(csr->row_ptr);
        host_free(csr->col);
    }
    csr->row_ptr = nullptr;
    csr->col = nullptr;
    csr->n_edges = 0;
    return IGC_OK;
}

// Device-to-host copy into pageable memory (d2h_stream with a plain move).
static int d2h_staged(igc_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    uint8_t *out = static_cast<uint8_t *>(dst);
    return igc::d2h_stream(ctx, src, bytes,
                      [out](const uint8_t *chunk, size_t off, size_t n) { std::memcpy(out + off, chunk, n); });
}

// Copies a device CSR of E edges (2E entries, or E for an upper-triangle CSR)
// into newly allocated host arrays, then applies the T-order of flags.
static int csr_to_host_locked(igc_ctx *ctx, uint32_t N, uint64_t E, const uint64_t *d_row_ptr, const uint32_t *d_col,
                              uint32_t flags, igc_csr *host)
{
    const bool upper = (flags & IGC_UPPER_TRIANGLE) != 0;
    const size_t ents = (upper ? 1 : 2) * (size_t)E;
    host->n_nodes = N;
    host->location = IGC_ON_HOST;
    host->n_edges = E;
    host->row_ptr = static_cast<uint64_t *>(host_alloc(sizeof(uint64_t) * ((size_t)N + 1)));
    host->col = static_cast<uint32_t *>(host_alloc(sizeof(uint32_t) * std::max<size_t>(ents, 1)));
    if (!host->row_ptr || !host->col) {
        igc_csr_free(host);
        return set_error(IGC_ERR_OOM, "host CSR allocation failed");
    }
    if (N == 0) {
        host->row_ptr[0] = 0;
        return IGC_OK;
    }
    int rc = d2h_staged(ctx, host->row_ptr, d_row_ptr, sizeof(uint64_t) * ((size_t)N + 1));
    if (!rc && E) rc = d2h_staged(ctx, host->col, d_col, sizeof(uint32_t) * ents);
    if (rc) {
        igc_csr_free(host);
        return rc;
    }
    const uint32_t T = (flags >> 16) & 0xFFFFu;
    if (T > 1 && !upper) igc::reorder_threads(*host, T); // upper rows are ascending in T-order too
    return IGC_OK;
}

IGC_API int igc_csr_to_host(igc_ctx *ctx, const igc_csr *dev, uint32_t flags, igc_csr *host)
{
    if (!ctx || !dev || !host) return set_error(IGC_ERR_INVALID, "null argument");
    if (dev->location != IGC_ON_DEVICE) return set_error(IGC_ERR_INVALID, "source CSR is not on the device");
    if (flags & IGC_UPPER_TRIANGLE) return set_error(IGC_ERR_INVALID, "device CSRs are symmetric; build with the flag");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return csr_to_host_locked(ctx, dev->n_nodes, dev->n_edges, dev->row_ptr, dev->col, flags, host);
}

IGC_API int igc_graph_build(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                            uint32_t flags, igc_csr *out)
{
    int rc = check_args(ctx, N, n);
    if (rc) return rc;
    if (!out || (!syms && N)) return set_error(IGC_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof *out);
    const bool on_device = (flags & IGC_RESULT_ON_DEVICE) != 0;
    if (on_device && ((flags >> 16) & 0xFFFFu) > 1)
        return set_error(IGC_ERR_INVALID, "a thread order needs the host CSR (drop IGC_RESULT_ON_DEVICE)");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t bytes = (size_t)N * n;
    if ((rc = ctx->syms.ensure(std::max<size_t>(bytes, 1)))) return rc;
    if (bytes) HIP_TRY(hipMemcpyAsync(ctx->syms.p, syms, bytes, hipMemcpyHostToDevice, ctx->stream));
    uint64_t E = 0;
    rc = build_device_locked(ctx, ctx->syms.as<uint8_t>(), N, n, edit_dist, &E, false,
                             (flags & IGC_UPPER_TRIANGLE) != 0);
    if (rc) return rc;
    if (on_device) {
        out->n_nodes = N;
        out->location = IGC_ON_DEVICE;
        out->n_edges = E;
        out->row_ptr = ctx->row_ptr.as<uint64_t>();
        out->col = ctx->col.as<uint32_t>();
        return IGC_OK;
    }
    return csr_to_host_locked(ctx, N, E, ctx->row_ptr.as<uint64_t>(), ctx->col.as<uint32_t>(), flags, out);
}

IGC_API int igc_autotune(igc_ctx *ctx, uint32_t n, uint32_t edit_dist, uint32_t N, uint32_t reps,
                         igc_tune_entry *entries, uint32_t cap, uint32_t *count)
{
    int rc = check_args(ctx, N, n);
    if (rc) return rc;
    if (!count || (!entries && cap)) return set_error(IGC_ERR_INVALID, "null argument");
    if (N < 2 || edit_dist == 0) return set_error(IGC_ERR_INVALID, "nothing to tune for N=%u editDist=%u", N, edit_dist);
    if (reps < 1) reps = 1;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int K = effective_k(n, edit_dist);
    if ((rc = ctx->packed.ensure(word_bytes(n) * (size_t)N))) return rc;
    if ((rc = ctx->misc.ensure(sizeof(uint32_t) * kMiscWords))) return rc;
    if (binary_words(n))
        hipLaunchKernelGGL((k_random_packed<uint64_t, true>), dim3((N + 255) / 256), dim3(256), 0, s,
                           ctx->packed.as<uint64_t>(), N, n, 0x5EEDull + N);
    else if (wide_words(n))
        hipLaunchKernelGGL(k_random_packed<uint64_t>, dim3((N + 255) / 256), dim3(256), 0, s,
                           ctx->packed.as<uint64_t>(), N, n, 0x5EEDull + N);
    else
        hipLaunchKernelGGL(k_random_packed<uint32_t>, dim3((N + 255) / 256), dim3(256), 0, s,
                           ctx->packed.as<uint32_t>(), N, n, 0x5EEDull + N);
    HIP_TRY(hipGetLastError());
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(ctx->misc.as<uint32_t>());
    struct Timed {
        Shape sh;
        float ms;
    };
    std::vector<Timed> timed;
    uint32_t k = 0;
    for (int cpl : {1, 2, 4}) {
        for (uint32_t q : {1u, 2u, 4u, 8u}) {
            const TileGeom G = choose_geometry(N, cpl * lane_cols(n), q);
            const uint64_t tiles = tile_count(G);
            if (tiles == 0 || tiles > 0x7FFFFFFFull) continue;
            float t_min = 0;
            for (uint32_t r = 0; r <= reps; ++r) { // first run warms up
                HIP_TRY(hipEventRecord(ctx->ev[0], s));
                if ((rc = launch_pairs(ctx_pred(ctx, n, K), n, K, cpl, false, s, tiles, ctx->packed.p, G, 0u,
                                       nullptr, 0, counter)))
                    return rc;
                HIP_TRY(hipEventRecord(ctx->ev[1], s));
                HIP_TRY(hipEventSynchronize(ctx->ev[1]));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
                if (r > 0 && (r == 1 || ms < t_min)) t_min = ms;
            }
            if (k < cap) entries[k] = igc_tune_entry{cpl, q, t_min, (float)((double)N * (N - 1) / 2 / (t_min * 1e6))};
            ++k;
            timed.push_back({Shape{cpl, q}, t_min});
        }
    }
    // the first shape (fewest slots, then fewest row blocks per group) within
    // 2% of the fastest: shapes that close differ by less than run-to-run noise
    Shape best_shape = default_shape(N);
    float best = 0;
    for (const Timed &t : timed)
        if (best == 0 || t.ms < best) best = t.ms;
    for (const Timed &t : timed) {
        if (t.ms <= best * 1.02f) {
            best_shape = t.sh;
            break;
        }
    }
    ctx->tuned[pred_variant(ctx_pred(ctx, n, K), n, K)][size_bucket(N)] = best_shape;
    *count = k;
    return IGC_OK;
}

IGC_API int igc_tuned_shape(igc_ctx *ctx, uint32_t n, uint32_t edit_dist, uint32_t N, int *cpl, uint32_t *q)
{
    if (!ctx || !cpl || !q) return set_error(IGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const Shape sh = pick_shape(ctx, n, effective_k(n, edit_dist), N);
    *cpl = sh.cpl;
    *q = sh.q;
    return IGC_OK;
}

IGC_API int igc_pair_plan_of(igc_ctx *ctx, uint32_t n, uint32_t edit_dist, uint32_t N, igc_pair_plan *out)
{
    int rc = check_args(ctx, N, n);
    if (rc) return rc;
    if (!out) return set_error(IGC_ERR_INVALID, "null argument");
    if (edit_dist == 0) return set_error(IGC_ERR_INVALID, "editDist 0 launches no pair kernel");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const int K = effective_k(n, edit_dist);
    const bool filtered = filter_pays(ctx->prefilter, N, n, K);
    const Shape sh = build_shape(ctx, n, K, N, filtered);
    const TileGeom G = choose_geometry(N, sh.cpl * lane_cols(n), sh.q);
    const PredKind pk = ctx_pred(ctx, n, K);
    out->predicate = pk == PredKind::LV ? IGC_PRED_LANDAU_VISHKIN : (pk == PredKind::MYERS ? IGC_PRED_MYERS : IGC_PRED_ALL_PAIRS);
    out->k = (int)edit_dist - 1;
    out->lane_cols = lane_cols(n);
    out->slots = sh.cpl;
    out->q = sh.q;
    out->tile_cols = G.W;
    out->tile_rows = G.R;
    out->tiles = tile_count(G);
    out->row_filter = filtered ? 1 : 0;
    out->pad = 0;
    return IGC_OK;
}

IGC_API int igc_graph_count(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                            uint64_t *n_edges)
{
    int rc = check_args(ctx, N, n);
    if (rc) return rc;
    if (!n_edges || (!syms && N)) return set_error(IGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t bytes = (size_t)N * n;
    if ((rc = ctx->syms.ensure(std::max<size_t>(bytes, 1)))) return rc;
    if (bytes) HIP_TRY(hipMemcpyAsync(ctx->syms.p, syms, bytes, hipMemcpyHostToDevice, ctx->stream));
    return build_device_locked(ctx, ctx->syms.as<uint8_t>(), N, n, edit_dist, n_edges, true);
}

} // extern "C"
