// igc_gen.hip -- candidate generators on the device (gfx950).
//
// The reference's LinearCode, VTCode, DifferentialVTCode and AllStrings
// generators each enumerate words in a fixed order, keep those that pass a
// test, and then drop the ones that fail applyFilters (max run of one symbol,
// GC fraction; src/CandidateGenerator.cpp:24-50).  Here one thread tests
// kGenItems consecutive words held as two symbol bit planes (bit p of `lo` /
// `hi` = low / high bit of symbol p), every test and filter is a handful of
// bitwise ops and popcounts on the planes, and the kept words are compacted in
// enumeration order in two passes: count per workgroup, scan, then recount and
// write (the same order as the reference's list, so the candidate ids, and the
// codebook the greedy picks from them, are the reference's).
//
//   LinearCode  message t (k base-4 digits, most significant first, the order
//               of NextBase4, src/Utils.cpp:525-545), codeword m*G + bias over
//               GF(4) (src/Candidates/LinearCodes.cpp:204-269).  GF(4)
//               addition is XOR of the 2-bit labels and multiplication by a
//               constant is linear over GF(2), so the codeword is the bias
//               XOR one precomputed row v*G[i] per message digit; a thread's
//               16 consecutive messages differ only in their last two digits.
//   VTCode      word t (base-4 index, last symbol least significant;
//               src/Candidates/VTCodes.cpp): kept when sum_{p>=1} p*[w_p >=
//               w_{p-1}] == a (mod n) and sum w_p == b (mod 4).
//   DifferentialVTCode  y_p = (w_p - w_{p+1}) mod 4, y_{n-1} = w_{n-1}; kept
//               when sum (p+1)*y_p == s (mod 4n) (src/Candidates/
//               DifferentialVTCodes.cpp).
//   AllStrings  every word.
// Weighted sums sum_p p*x_p of a bit mask x are five popcounts of x against the
// masks of positions whose bit b is set.
#include "igc_device.hpp"

#include <cmath>
#include <vector>

namespace igc {

constexpr int kGenThreads = 256; // = kScanThreads: the write pass uses block_exclusive_scan
constexpr int kGenItems = 16;    // consecutive words per thread (the last two base-4 digits)
constexpr uint64_t kGenBlock = (uint64_t)kGenThreads * kGenItems;
static_assert(kGenThreads == kScanThreads, "block scan width");

enum GenKind : int { GEN_LINEAR = 0, GEN_VT = 1, GEN_DVT = 2, GEN_ALL = 3 };

struct GenParams {
    uint64_t total;   // words enumerated
    uint64_t gc_ok;   // bit c: a string with c symbols in {1, 2} passes the GC filter
    const uint2 *rows;// LinearCode: rows[4i + v] = planes of v * G[i]
    uint32_t n;       // string length
    uint32_t k;       // LinearCode message length
    uint32_t max_run; // longest allowed run (0: no run filter)
    uint32_t a;       // VT: alpha target mod n; DVT: syndrome target mod 4n
    uint32_t b;       // VT: symbol-sum target mod 4
    uint32_t bias_lo, bias_hi;
    uint32_t pad;
};

__device__ __forceinline__ uint32_t weighted_popc(uint32_t x) // sum of p over set bits p
{
    return __popc(x & 0xAAAAAAAAu) + (__popc(x & 0xCCCCCCCCu) << 1) + (__popc(x & 0xF0F0F0F0u) << 2) +
           (__popc(x & 0xFF00FF00u) << 3) + (__popc(x & 0xFFFF0000u) << 4);
}

// bit j of the result = bit 2j of x
__device__ __forceinline__ uint32_t even_bits(uint64_t x)
{
    x &= 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
    return (uint32_t)x;
}

// Planes of word t of the base-4 enumeration: symbol p is digit n-1-p of t.
__device__ __forceinline__ void index_planes(uint64_t t, uint32_t n, uint32_t &lo, uint32_t &hi)
{
    lo = __builtin_bitreverse32(even_bits(t)) >> (32 - n);
    hi = __builtin_bitreverse32(even_bits(t >> 1)) >> (32 - n);
}

template <int KIND>
__device__ __forceinline__ bool gen_keep(const GenParams &P, uint32_t lo, uint32_t hi)
{
    const uint32_t n = P.n;
    const uint32_t mask = n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
    const uint32_t inner = mask & ~1u; // positions 1..n-1
    const uint32_t lo1 = lo << 1, hi1 = hi << 1; // bit p: symbol p-1
    if constexpr (KIND == GEN_VT) {
        const uint32_t ge = ((hi & ~hi1) | (~(hi ^ hi1) & (lo | ~lo1))) & inner;
        if (weighted_popc(ge) % n != P.a) return false;
        if (((__popc(lo) + 2u * __popc(hi)) & 3u) != P.b) return false;
    } else if constexpr (KIND == GEN_DVT) {
        const uint32_t lon = lo >> 1, hin = hi >> 1; // bit p: symbol p+1 (0 past the end)
        const uint32_t ylo = lo ^ lon, yhi = (hi ^ hin ^ (~lo & lon)) & mask;
        const uint32_t syn = weighted_popc(ylo) + __popc(ylo) + 2u * (weighted_popc(yhi) + __popc(yhi));
        if (syn % (4u * n) != P.a) return false;
    }
    if (P.max_run) {
        const uint32_t eq = ~((lo ^ lo1) | (hi ^ hi1)) & inner; // bit p: symbol p equals symbol p-1
        uint32_t x = eq;
        for (uint32_t r = 1; r < P.max_run && x; ++r) x &= eq >> r;
        if (x) return false; // max_run equalities in a row: a run of max_run + 1
    }
    return (P.gc_ok >> __popc((lo ^ hi) & mask)) & 1ull;
}

// Planes of the kGenItems words of this thread and which of them are kept.
template <int KIND>
__device__ __forceinline__ uint32_t gen_thread(const GenParams &P, const uint2 *rows_s, uint64_t t0, uint32_t *lo,
                                               uint32_t *hi)
{
    uint32_t keep = 0;
    if (t0 >= P.total) return 0;
    if constexpr (KIND == GEN_LINEAR) {
        // digits above the last two are fixed for the thread's 16 messages
        uint32_t blo = P.bias_lo, bhi = P.bias_hi;
        const int k = (int)P.k;
        for (int i = 0; i + 2 < k; ++i) {
            const uint32_t d = (uint32_t)(t0 >> (2 * (k - 1 - i))) & 3u;
            blo ^= rows_s[4 * i + d].x;
            bhi ^= rows_s[4 * i + d].y;
        }
#pragma unroll
        for (int j = 0; j < kGenItems; ++j) {
            uint32_t wl = blo, wh = bhi;
            wl ^= rows_s[4 * (k - 1) + (j & 3)].x;
            wh ^= rows_s[4 * (k - 1) + (j & 3)].y;
            if (k >= 2) {
                wl ^= rows_s[4 * (k - 2) + (j >> 2)].x;
                wh ^= rows_s[4 * (k - 2) + (j >> 2)].y;
            }
            lo[j] = wl;
            hi[j] = wh;
            if (t0 + j < P.total && gen_keep<KIND>(P, wl, wh)) keep |= 1u << j;
        }
    } else {
        // t0 is a multiple of 16: word t0 + j is word t0 with its last two
        // symbols (digits 0 and 1 of the index) set from j
        uint32_t l0, h0;
        index_planes(t0, P.n, l0, h0);
        const uint32_t p1 = P.n - 1, p2 = P.n >= 2 ? P.n - 2 : 0;
#pragma unroll
        for (int j = 0; j < kGenItems; ++j) {
            lo[j] = l0 | ((uint32_t)(j & 1) << p1) | (P.n >= 2 ? (uint32_t)((j >> 2) & 1) << p2 : 0u);
            hi[j] = h0 | ((uint32_t)((j >> 1) & 1) << p1) | (P.n >= 2 ? (uint32_t)((j >> 3) & 1) << p2 : 0u);
            if (t0 + j < P.total && gen_keep<KIND>(P, lo[j], hi[j])) keep |= 1u << j;
        }
    }
    return keep;
}

__device__ __forceinline__ void load_rows(const GenParams &P, uint2 *rows_s)
{
    for (uint32_t i = threadIdx.x; i < 4 * P.k; i += blockDim.x) rows_s[i] = P.rows[i];
    __syncthreads();
}

constexpr int kGenMaxK = 15; // 4^15 messages

template <int KIND>
__global__ __launch_bounds__(kGenThreads) void k_gen_count(GenParams P, uint32_t *__restrict__ blk_count)
{
    __shared__ uint2 rows_s[4 * kGenMaxK];
    __shared__ uint32_t wsum[kGenThreads / 64];
    if constexpr (KIND == GEN_LINEAR) load_rows(P, rows_s);
    uint32_t lo[kGenItems], hi[kGenItems];
    const uint64_t t0 = (uint64_t)blockIdx.x * kGenBlock + (uint64_t)threadIdx.x * kGenItems;
    uint32_t c = __popc(gen_thread<KIND>(P, rows_s, t0, lo, hi));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < kGenThreads / 64; ++w) s += wsum[w];
        blk_count[blockIdx.x] = s;
    }
}

template <int KIND>
__global__ __launch_bounds__(kGenThreads) void k_gen_write(GenParams P, const uint64_t *__restrict__ blk_off,
                                                           uint8_t *__restrict__ out)
{
    __shared__ uint2 rows_s[4 * kGenMaxK];
    __shared__ uint64_t sh[kScanThreads / 64 + 1];
    __shared__ uint2 staged[kGenThreads / 64][64 * kGenItems]; // a wave's kept words, in output order
    if constexpr (KIND == GEN_LINEAR) load_rows(P, rows_s);
    uint32_t lo[kGenItems], hi[kGenItems];
    const uint64_t t0 = (uint64_t)blockIdx.x * kGenBlock + (uint64_t)threadIdx.x * kGenItems;
    const uint32_t keep = gen_thread<KIND>(P, rows_s, t0, lo, hi);
    uint64_t excl;
    block_exclusive_scan((uint64_t)__popc(keep), sh, excl);
    const uint64_t o = blk_off[blockIdx.x] + excl;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the wave's kept words are output words wo .. wo + wc - 1
    const uint64_t wo = uniform64(o);
    const uint64_t end = o + __popc(keep);
    const uint64_t wend = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(end >> 32), 63, 64) << 32) |
                          (uint32_t)__shfl((int)(uint32_t)end, 63, 64);
    const uint32_t wc = (uint32_t)(wend - wo);
    uint32_t r = (uint32_t)(o - wo);
#pragma unroll
    for (int j = 0; j < kGenItems; ++j)
        if ((keep >> j) & 1u) staged[wave][r++] = make_uint2(lo[j], hi[j]);
    __syncthreads();
    // bytes [b0, b1) of the output, written as aligned 32-bit words by
    // consecutive lanes; the (at most 3 + 3) bytes outside the aligned part
    // are byte stores (the neighbouring wave writes the rest of those words).
    const uint32_t n = P.n;
    const uint32_t inv = ((1u << 20) + n - 1) / n; // r / n = (r * inv) >> 20 for r < 2^15, n <= 32
    const uint64_t b0 = wo * n, b1 = b0 + (uint64_t)wc * n;
    const uint64_t a0 = (b0 + 3) & ~3ull, a1 = b1 & ~3ull;
    const uint2 *st = staged[wave];
    auto sym_at = [&](uint64_t pos) -> uint32_t {
        const uint32_t rr = (uint32_t)(pos - b0);
        const uint32_t w = (rr * inv) >> 20, p = rr - w * n;
        const uint2 pl = st[w];
        return ((pl.x >> p) & 1u) | (((pl.y >> p) & 1u) << 1);
    };
    if (a0 < a1) {
        for (uint64_t q = a0 + 4ull * lane; q < a1; q += 256)
            *reinterpret_cast<uint32_t *>(out + q) =
                sym_at(q) | (sym_at(q + 1) << 8) | (sym_at(q + 2) << 16) | (sym_at(q + 3) << 24);
        if (lane < a0 - b0) out[b0 + lane] = (uint8_t)sym_at(b0 + lane);
        if (lane >= 32 && lane - 32 < b1 - a1) out[a1 + lane - 32] = (uint8_t)sym_at(a1 + lane - 32);
    } else if (lane < b1 - b0) {
        out[b0 + lane] = (uint8_t)sym_at(b0 + lane);
    }
}

// GC counts c in 0..n that pass [min_gc, max_gc], with the reference's double
// arithmetic (GCContent = double(c) / double(n), src/Utils.cpp:447-462).
static uint64_t gc_mask(uint32_t n, double min_gc, double max_gc)
{
    if (!(min_gc > 0.0 || max_gc > 0.0)) return ~0ull;
    uint64_t m = 0;
    for (uint32_t c = 0; c <= n; ++c) {
        const double f = double(c) / double(n);
        if (f >= min_gc && f <= max_gc) m |= 1ull << c;
    }
    return m;
}

template <int KIND>
static void launch_gen(bool write, hipStream_t s, uint32_t blocks, const GenParams &P, uint32_t *cnt,
                       const uint64_t *off, uint8_t *out)
{
    if (write)
        hipLaunchKernelGGL(k_gen_write<KIND>, dim3(blocks), dim3(kGenThreads), 0, s, P, off, out);
    else
        hipLaunchKernelGGL(k_gen_count<KIND>, dim3(blocks), dim3(kGenThreads), 0, s, P, cnt);
}

static void launch_gen_kind(int kind, bool write, hipStream_t s, uint32_t blocks, const GenParams &P, uint32_t *cnt,
                            const uint64_t *off, uint8_t *out)
{
    switch (kind) {
    case GEN_LINEAR: launch_gen<GEN_LINEAR>(write, s, blocks, P, cnt, off, out); break;
    case GEN_VT: launch_gen<GEN_VT>(write, s, blocks, P, cnt, off, out); break;
    case GEN_DVT: launch_gen<GEN_DVT>(write, s, blocks, P, cnt, off, out); break;
    default: launch_gen<GEN_ALL>(write, s, blocks, P, cnt, off, out); break;
    }
}

static int gen_run(igc_ctx *ctx, int kind, GenParams P, const igc_filter *filter, igc_candidates *out)
{
    if (filter) {
        P.max_run = filter->max_run >= P.n ? 0u : filter->max_run;
        P.gc_ok = gc_mask(P.n, filter->min_gc, filter->max_gc);
    } else {
        P.max_run = 0;
        P.gc_ok = ~0ull;
    }
    const uint64_t blocks64 = (P.total + kGenBlock - 1) / kGenBlock;
    if (blocks64 > 0x7FFFFFFFull) return set_error(IGC_ERR_INVALID, "enumeration of %llu words is too large",
                                                   (unsigned long long)P.total);
    const uint32_t blocks = (uint32_t)blocks64;
    hipStream_t s = ctx->stream;
    int rc;
    const uint32_t nsb = (uint32_t)(((uint64_t)blocks + kScanTile - 1) / kScanTile);
    if ((rc = ctx->gen_cnt.ensure(sizeof(uint32_t) * (size_t)blocks))) return rc;
    if ((rc = ctx->gen_off.ensure(sizeof(uint64_t) * ((size_t)blocks + 1)))) return rc;
    if ((rc = ctx->bsum.ensure(sizeof(uint64_t) * std::max<size_t>(nsb, 1)))) return rc;
    HIP_TRY(hipEventRecord(ctx->ev[0], s));
    launch_gen_kind(kind, false, s, blocks, P, ctx->gen_cnt.as<uint32_t>(), nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    scan_counts(s, ctx->gen_cnt.as<uint32_t>(), blocks, ctx->bsum.as<uint64_t>(), ctx->gen_off.as<uint64_t>());
    HIP_TRY(hipGetLastError());
    uint64_t kept = 0;
    HIP_TRY(hipMemcpyAsync(&kept, ctx->gen_off.as<uint64_t>() + blocks, sizeof kept, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if ((rc = ctx->gen_out.ensure(std::max<size_t>((size_t)kept * P.n, 1)))) return rc;
    launch_gen_kind(kind, true, s, blocks, P, nullptr, ctx->gen_off.as<uint64_t>(), ctx->gen_out.as<uint8_t>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[1], s));
    HIP_TRY(hipEventSynchronize(ctx->ev[1]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    out->d_syms = ctx->gen_out.as<uint8_t>();
    out->serial = ++ctx->gen_serial;
    out->count = kept;
    out->enumerated = P.total;
    out->n = P.n;
    out->ms = ms;
    return IGC_OK;
}

static int gen_args(igc_ctx *ctx, uint32_t n, uint32_t nmax, const igc_filter *filter, igc_candidates *out)
{
    if (!ctx || !out) return set_error(IGC_ERR_INVALID, "null argument");
    if (n < 1 || n > nmax) return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%u", n, nmax);
    if (filter && (std::isnan(filter->min_gc) || std::isnan(filter->max_gc)))
        return set_error(IGC_ERR_INVALID, "GC bounds must be numbers");
    return IGC_OK;
}

} // namespace igc

using namespace igc;

extern "C" {

IGC_API int igc_gen_linear(igc_ctx *ctx, const uint8_t *gen, uint32_t k, uint32_t n, const uint8_t *bias,
                           const igc_filter *filter, igc_candidates *out)
{
    int rc = gen_args(ctx, n, IGC_MAX_QUATERNARY_LEN, filter, out);
    if (rc) return rc;
    if (!gen) return set_error(IGC_ERR_INVALID, "null generator matrix");
    if (k < 1 || k > (uint32_t)kGenMaxK) return set_error(IGC_ERR_INVALID, "message length k=%u outside 1..%d", k, kGenMaxK);
    // rows[4i + v] = planes of v * G[i]; GF(4) with x^2 + x + 1, labels 0, 1, x, x + 1
    static const uint8_t mul[4][4] = {{0, 0, 0, 0}, {0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}};
    std::vector<uint2> rows(4 * (size_t)k);
    for (uint32_t i = 0; i < k; ++i)
        for (uint32_t v = 0; v < 4; ++v) {
            uint32_t lo = 0, hi = 0;
            for (uint32_t j = 0; j < n; ++j) {
                const uint8_t g = gen[(size_t)i * n + j];
                if (g > 3) return set_error(IGC_ERR_INVALID, "generator entry %u at (%u, %u) is not in GF(4)", g, i, j);
                const uint8_t c = mul[v][g];
                lo |= (uint32_t)(c & 1u) << j;
                hi |= (uint32_t)(c >> 1) << j;
            }
            rows[4 * i + v] = make_uint2(lo, hi);
        }
    GenParams P{};
    P.total = 1ull << (2 * k);
    P.n = n;
    P.k = k;
    if (bias)
        for (uint32_t j = 0; j < n; ++j) {
            if (bias[j] > 3) return set_error(IGC_ERR_INVALID, "bias entry %u at %u is not in GF(4)", bias[j], j);
            P.bias_lo |= (uint32_t)(bias[j] & 1u) << j;
            P.bias_hi |= (uint32_t)(bias[j] >> 1) << j;
        }
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    if ((rc = ctx->gen_rows.ensure(sizeof(uint2) * rows.size()))) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->gen_rows.p, rows.data(), sizeof(uint2) * rows.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    P.rows = ctx->gen_rows.as<uint2>();
    return gen_run(ctx, GEN_LINEAR, P, filter, out);
}

IGC_API int igc_gen_vt(igc_ctx *ctx, uint32_t n, int32_t a, int32_t b, const igc_filter *filter, igc_candidates *out)
{
    int rc = gen_args(ctx, n, IGC_GEN_MAX_WORD_LEN, filter, out);
    if (rc) return rc;
    GenParams P{};
    P.total = 1ull << (2 * n);
    P.n = n;
    P.a = (uint32_t)((a % (int32_t)n + (int32_t)n) % (int32_t)n);
    P.b = (uint32_t)((b % 4 + 4) % 4);
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return gen_run(ctx, GEN_VT, P, filter, out);
}

IGC_API int igc_gen_diff_vt(igc_ctx *ctx, uint32_t n, int32_t syndrome, const igc_filter *filter, igc_candidates *out)
{
    int rc = gen_args(ctx, n, IGC_GEN_MAX_WORD_LEN, filter, out);
    if (rc) return rc;
    GenParams P{};
    P.total = 1ull << (2 * n);
    P.n = n;
    const int64_t mod = 4 * (int64_t)n;
    P.a = (uint32_t)(((int64_t)syndrome % mod + mod) % mod);
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return gen_run(ctx, GEN_DVT, P, filter, out);
}

IGC_API int igc_gen_all(igc_ctx *ctx, uint32_t n, const igc_filter *filter, igc_candidates *out)
{
    int rc = gen_args(ctx, n, IGC_GEN_MAX_WORD_LEN, filter, out);
    if (rc) return rc;
    GenParams P{};
    P.total = 1ull << (2 * n);
    P.n = n;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return gen_run(ctx, GEN_ALL, P, filter, out);
}

IGC_API int igc_candidates_to_host(igc_ctx *ctx, const igc_candidates *c, uint8_t *out, uint64_t cap)
{
    if (!ctx || !c || (!out && cap)) return set_error(IGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (c->serial != ctx->gen_serial || c->d_syms != ctx->gen_out.as<uint8_t>())
        return set_error(IGC_ERR_INVALID, "candidates are not this context's latest generation");
    const uint64_t m = std::min<uint64_t>(cap, c->count);
    if (m == 0) return IGC_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMemcpyAsync(out, c->d_syms, (size_t)m * c->n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return IGC_OK;
}

} // extern "C"
