// igc_prefilter.hpp -- exact row skipping for the pair kernel (gfx950).
//
// Bound.  Split every string at h = n/2 and count its symbols in each half:
// p(A) and s(A), four counts each.  For equal-length A and B,
//   |p(A) - p(B)|_1 + |s(A) - s(B)|_1 <= 2 * Levenshtein(A, B).
// A substitution changes one count pair by 1 each (2).  Because the lengths
// are equal, insertions and deletions come in pairs; applying one of each
// shifts the segment between them by one position, so besides the inserted
// and the deleted symbol at most one symbol crosses the split: 4 per pair, 2
// per edit.  So an edge (Levenshtein <= K) needs that L1 distance <= 2K.
//
// Order.  Candidates are bucket-sorted by their count class (p, s), the
// classes taken along a Hilbert curve through (p0, p1, p2, s0, s1, s2) (the
// fourth count of each half follows from the other three), so the 64*slots*P
// columns a wave owns (a "window") hold a few neighbouring classes.
//
// Items.  Each window keeps up to kItemCap items: one per run of equal
// classes (a point), or, when a window holds more runs, kItemCap boxes of
// consecutive runs (byte-wise min and max).  For a row r and an item [lo, hi]
//   sum_d |r_d - lo_d| + |r_d - hi_d| = sum_d (hi_d - lo_d) + 2 * dist(r, item),
// i.e. four v_sad_u8 on the packed counts started from a per-item bias that
// cancels the width.  The row is skipped for the whole window when every
// item is farther than 2K.  Skipping is exact: no pair that can be an edge is
// skipped, so the graph is the same; edges are emitted with the original ids
// through the permutation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <utility>
#include <vector>

namespace igc {

constexpr uint32_t kItemCap = 16;          // items per window
constexpr uint32_t kItemBias = 1u << 16;   // above any width (8 counts of at most 32)

// Pair-kernel view of the prefilter (all null: no filtering, identity ids).
struct RowFilter {
    const uint2 *feat;    // counts of the sorted candidates: x = prefix, y = suffix (a byte per symbol)
    const uint4 *items;   // window w's items at w * kItemCap: x/y = byte-wise min prefix/suffix, z/w = max
    const uint32_t *bias; // per item: kItemBias - sum(hi - lo)
    const uint32_t *count;// items per window
    const uint32_t *perm; // sorted position -> original id
    uint32_t window;      // columns per window (a wave's column span)
    uint32_t limit;       // kItemBias + 4K: a row is kept when some item's SAD chain is at most this
};

__host__ __device__ inline RowFilter filter_view(const uint2 *feat, const uint4 *items, const uint32_t *bias,
                                                 const uint32_t *count, const uint32_t *perm, uint32_t window, int K)
{
    RowFilter f{};
    f.feat = feat;
    f.items = items;
    f.bias = bias;
    f.count = count;
    f.perm = perm;
    f.window = window;
    f.limit = kItemBias + 4u * (uint32_t)K;
    return f;
}

__device__ __forceinline__ uint32_t half_counts(const uint8_t *s, uint32_t a, uint32_t b)
{
    uint32_t c = 0;
    for (uint32_t p = a; p < b; ++p) c += 1u << (8 * (s[p] & 3u));
    return c;
}

// Rank of the composition (c0, c1, c2, c3) of `len` symbols among all such
// compositions in lexicographic order.
__device__ __forceinline__ uint32_t comp_rank(uint32_t packed, uint32_t len)
{
    const uint32_t c0 = packed & 0xFF, c1 = (packed >> 8) & 0xFF, c2 = (packed >> 16) & 0xFF;
    uint32_t r = 0;
    for (uint32_t a = 0; a < c0; ++a) {
        const uint32_t m = len - a; // the other three symbols share m
        r += (m + 2) * (m + 1) / 2;
    }
    for (uint32_t b = 0; b < c1; ++b) r += len - c0 - b + 1;
    return r + c2;
}

__host__ __device__ inline uint32_t comp_count(uint32_t len) { return (len + 3) * (len + 2) * (len + 1) / 6; }

// One set of a filter set-up; several sets (a batch) share one set of launches.
struct FilterSet {
    uint64_t sym_off;      // the set's symbols (byte offset)
    uint64_t word_off;     // the set's packed words (byte offset; 64-bit words for n > 16)
    const uint32_t *order; // class order of the set's string length
    uint32_t node_off;     // first node of the set (sorted counts and permutation are per node)
    uint32_t N, n;
    uint32_t fbase;        // first position of the set among all sets' strings
    uint32_t kbase;        // first bucket of the set
    uint32_t wbase;        // first window of the set
    uint32_t window;       // columns per window
    uint32_t pad;
};

// Set holding position `pos` of the field (fbase or wbase) `by_window` picks.
__device__ __forceinline__ uint32_t filter_set_of(const FilterSet *__restrict__ fs, uint32_t nf, uint32_t pos,
                                                  bool by_window)
{
    uint32_t lo = 0, hi = nf - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if ((by_window ? fs[mid].wbase : fs[mid].fbase) <= pos) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// For string g of all sets: its half counts feat[g], its bucket key[g] = the
// set's first bucket + the position of its count class along the curve
// (order[rank(prefix) * comp_count(n-h) + rank(suffix)]), and the bucket
// histogram.
static __global__ void k_fs_keys(const uint8_t *__restrict__ syms, const FilterSet *__restrict__ fs, uint32_t nf,
                                 uint32_t total, uint2 *__restrict__ feat, uint32_t *__restrict__ key,
                                 uint32_t *__restrict__ hist)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const FilterSet &f = fs[filter_set_of(fs, nf, g, false)];
    const uint32_t n = f.n, h = n / 2;
    const uint8_t *s = syms + f.sym_off + (size_t)(g - f.fbase) * n;
    const uint32_t p = half_counts(s, 0, h), q = half_counts(s, h, n);
    feat[g] = make_uint2(p, q);
    const uint32_t k = f.kbase + f.order[comp_rank(p, h) * comp_count(n - h) + comp_rank(q, n - h)];
    key[g] = k;
    atomicAdd(&hist[k], 1u);
}

// Counting-sort scatter: positions of a bucket are handed out from its end
// (any order inside a bucket will do).  A set's buckets follow the previous
// sets', so its sorted positions are fbase .. fbase+N-1; perm[node_off + k] =
// the set-local string at sorted position k.
static __global__ void k_fs_scatter(const uint32_t *__restrict__ key, const FilterSet *__restrict__ fs, uint32_t nf,
                                    uint32_t total, const uint64_t *__restrict__ off, uint32_t *__restrict__ hist,
                                    uint32_t *__restrict__ perm)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const FilterSet &f = fs[filter_set_of(fs, nf, g, false)];
    const uint32_t k = key[g];
    const uint32_t r = atomicSub(&hist[k], 1u) - 1u;
    perm[f.node_off + (uint32_t)(off[k] + r - f.fbase)] = g - f.fbase;
}

// Sorted words and counts of every set, position by position.
static __global__ void k_fs_permute(const FilterSet *__restrict__ fs, uint32_t nf, uint32_t total,
                                    const uint32_t *__restrict__ perm, const uint8_t *__restrict__ packed,
                                    const uint2 *__restrict__ feat, uint8_t *__restrict__ packed_s,
                                    uint2 *__restrict__ feat_s)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const FilterSet &f = fs[filter_set_of(fs, nf, g, false)];
    const uint32_t k = g - f.fbase;
    const uint32_t i = perm[f.node_off + k];
    if (f.n > 16)
        reinterpret_cast<uint64_t *>(packed_s + f.word_off)[k] =
            reinterpret_cast<const uint64_t *>(packed + f.word_off)[i];
    else
        reinterpret_cast<uint32_t *>(packed_s + f.word_off)[k] =
            reinterpret_cast<const uint32_t *>(packed + f.word_off)[i];
    feat_s[f.node_off + k] = feat[f.fbase + i];
}

__device__ __forceinline__ uint32_t byte_of(const uint2 f, uint32_t b)
{
    return ((b < 4 ? f.x : f.y) >> (8 * (b & 3))) & 0xFFu;
}

// Items of every window of every set (a set's windows hold `window` of its
// sorted candidates): one workgroup of 64 per window.  A column starts a run when its class differs from the column
// before it; run r goes to item r (at most kItemCap runs) or to item
// r * kItemCap / runs, and each item's byte-wise min and max are gathered in
// LDS.
static __global__ __launch_bounds__(64) void k_fs_items(const FilterSet *__restrict__ fs, uint32_t nf,
                                                        const uint2 *__restrict__ feat_s_all,
                                                        uint4 *__restrict__ items, uint32_t *__restrict__ bias,
                                                        uint32_t *__restrict__ count)
{
    __shared__ uint32_t lo[kItemCap * 8], hi[kItemCap * 8];
    const uint32_t w = blockIdx.x, lane = threadIdx.x; // w: window of all sets
    const FilterSet &f = fs[filter_set_of(fs, nf, w, true)];
    const uint2 *__restrict__ feat_s = feat_s_all + f.node_off;
    const uint32_t N = f.N, window = f.window;
    const uint32_t first = (w - f.wbase) * window;
    if (first >= N) return;
    const uint32_t last = min(first + window, N);
    for (uint32_t t = lane; t < kItemCap * 8; t += 64) {
        lo[t] = 0xFFu;
        hi[t] = 0u;
    }
    uint32_t runs = 0;
    for (uint32_t c0 = first; c0 < last; c0 += 64) {
        const uint32_t c = c0 + lane;
        bool start = false;
        if (c < last) {
            const uint2 f = feat_s[c];
            start = c == first || f.x != feat_s[c - 1].x || f.y != feat_s[c - 1].y;
        }
        runs += (uint32_t)__popcll(__ballot(start));
    }
    __syncthreads();
    uint32_t seen = 0;
    for (uint32_t c0 = first; c0 < last; c0 += 64) {
        const uint32_t c = c0 + lane;
        bool start = false;
        uint2 f = make_uint2(0u, 0u);
        if (c < last) {
            f = feat_s[c];
            start = c == first || f.x != feat_s[c - 1].x || f.y != feat_s[c - 1].y;
        }
        const uint64_t m = __ballot(start);
        if (c < last) {
            // runs started at or before this column, minus one
            const uint32_t run = seen + (uint32_t)__popcll(m & ((2ull << lane) - 1ull)) - 1u;
            const uint32_t it = runs <= kItemCap ? run : (uint32_t)((uint64_t)run * kItemCap / runs);
            for (uint32_t b = 0; b < 8; ++b) {
                atomicMin(&lo[it * 8 + b], byte_of(f, b));
                atomicMax(&hi[it * 8 + b], byte_of(f, b));
            }
        }
        seen += (uint32_t)__popcll(m);
    }
    __syncthreads();
    const uint32_t ni = min(runs, kItemCap);
    if (lane < ni) {
        uint32_t v[4] = {0u, 0u, 0u, 0u}, width = 0;
        for (uint32_t b = 0; b < 8; ++b) {
            const uint32_t l = lo[lane * 8 + b], u = hi[lane * 8 + b];
            v[b >> 2] |= l << (8 * (b & 3));
            v[2 + (b >> 2)] |= u << (8 * (b & 3));
            width += u - l;
        }
        items[(size_t)w * kItemCap + lane] = make_uint4(v[0], v[1], v[2], v[3]);
        bias[(size_t)w * kItemCap + lane] = kItemBias - width;
    }
    if (lane == 0) count[w] = ni;
}

// Position of every count class (rank(prefix) * comp_count(n-h) + rank(suffix))
// along a Hilbert curve through (p0, p1, p2, s0, s1, s2), 5 bits per axis.
// The curve index is J. Skilling's transpose form ("Programming the Hilbert
// curve", AIP Conf. Proc. 707, 2004): undo the excess work, Gray-encode, then
// interleave the axes' bits from the top.  Host code.
inline uint64_t hilbert_index6(uint32_t X[6])
{
    constexpr int D = 6, B = 5;
    const uint32_t M = 1u << (B - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < D; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < D; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[D - 1] & Q) t ^= Q - 1;
    for (int i = 0; i < D; ++i) X[i] ^= t;
    uint64_t key = 0;
    for (int b = B - 1; b >= 0; --b)
        for (int i = 0; i < D; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
    return key;
}

// Compositions of `len` into (c0, c1, c2, c3) in the lexicographic order
// comp_rank numbers them; three counts per composition.
inline void compositions(uint32_t len, std::vector<uint32_t> &out)
{
    out.clear();
    for (uint32_t a = 0; a <= len; ++a)
        for (uint32_t b = 0; a + b <= len; ++b)
            for (uint32_t c = 0; a + b + c <= len; ++c) {
                out.push_back(a);
                out.push_back(b);
                out.push_back(c);
            }
}

inline void class_order(uint32_t n, std::vector<uint32_t> &order)
{
    const uint32_t h = n / 2;
    std::vector<uint32_t> cp, cs;
    compositions(h, cp);
    compositions(n - h, cs);
    const uint32_t np = (uint32_t)cp.size() / 3, ns = (uint32_t)cs.size() / 3;
    std::vector<std::pair<uint64_t, uint32_t>> keyed((size_t)np * ns);
    for (uint32_t a = 0; a < np; ++a)
        for (uint32_t b = 0; b < ns; ++b) {
            uint32_t X[6] = {cp[3 * a], cp[3 * a + 1], cp[3 * a + 2], cs[3 * b], cs[3 * b + 1], cs[3 * b + 2]};
            keyed[(size_t)a * ns + b] = {hilbert_index6(X), a * ns + b};
        }
    std::sort(keyed.begin(), keyed.end());
    order.assign(keyed.size(), 0u);
    for (uint32_t r = 0; r < (uint32_t)keyed.size(); ++r) order[keyed[r].second] = r;
}

} // namespace igc
