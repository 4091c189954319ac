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
// Use.  Candidates are bucket-sorted by (prefix counts, suffix counts) in
// lexicographic order, so the 64*slots*P columns a wave owns (a "window")
// have similar counts.  Per window the byte-wise minimum and maximum of the
// eight counts form a box; for a row r the L1 distance to the box is a lower
// bound for its distance to every column in the window, and
//   sum_d |r_d - lo_d| + |r_d - hi_d| = sum_d (hi_d - lo_d) + 2 * dist(r, box),
// i.e. four v_sad_u8 on the packed counts.  The row is skipped for the whole
// window when dist > 2K.  Skipping is exact: no pair that can be an edge is
// skipped, so the graph is the same; edges are emitted with the original ids
// through the permutation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace igc {

// Pair-kernel view of the prefilter (all null: no filtering, identity ids).
struct RowFilter {
    const uint2 *feat;    // counts of the sorted candidates: x = prefix, y = suffix (a byte per symbol)
    const uint4 *box;     // per window: x/y = byte-wise min prefix/suffix, z/w = max
    const uint32_t *thr;  // per window: sum(hi - lo) + 4K (skip when the SAD sum exceeds it)
    const uint32_t *perm; // sorted position -> original id
    uint32_t window;      // columns per window (a wave's column span)
};

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

// feat[i], bucket key[i] = rank(prefix) * comp_count(n-h) + rank(suffix), and
// the bucket histogram.
static __global__ void k_node_keys(const uint8_t *__restrict__ syms, uint32_t N, uint32_t n,
                                   uint2 *__restrict__ feat, uint32_t *__restrict__ key,
                                   uint32_t *__restrict__ hist)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t h = n / 2;
    const uint8_t *s = syms + (size_t)i * n;
    const uint32_t p = half_counts(s, 0, h), q = half_counts(s, h, n);
    feat[i] = make_uint2(p, q);
    const uint32_t k = comp_rank(p, h) * comp_count(n - h) + comp_rank(q, n - h);
    key[i] = k;
    atomicAdd(&hist[k], 1u);
}

// Counting-sort scatter: positions of a bucket are handed out from its end
// (any order inside a bucket will do).
static __global__ void k_node_scatter(const uint32_t *__restrict__ key, uint32_t N, const uint64_t *__restrict__ off,
                                      uint32_t *__restrict__ hist, uint32_t *__restrict__ perm)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t k = key[i];
    const uint32_t r = atomicSub(&hist[k], 1u) - 1u;
    perm[off[k] + r] = i;
}

template <class W>
__global__ void k_node_permute(const uint32_t *__restrict__ perm, uint32_t N, const W *__restrict__ packed,
                               const uint2 *__restrict__ feat, W *__restrict__ packed_s, uint2 *__restrict__ feat_s)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const uint32_t i = perm[k];
    packed_s[k] = packed[i];
    feat_s[k] = feat[i];
}

__device__ __forceinline__ uint32_t bytes_min(uint32_t a, uint32_t b)
{
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) {
        const uint32_t x = (a >> (8 * k)) & 0xFF, y = (b >> (8 * k)) & 0xFF;
        r |= (x < y ? x : y) << (8 * k);
    }
    return r;
}

__device__ __forceinline__ uint32_t bytes_max(uint32_t a, uint32_t b)
{
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) {
        const uint32_t x = (a >> (8 * k)) & 0xFF, y = (b >> (8 * k)) & 0xFF;
        r |= (x > y ? x : y) << (8 * k);
    }
    return r;
}

__device__ __forceinline__ uint32_t bytes_sum(uint32_t a)
{
    return (a & 0xFF) + ((a >> 8) & 0xFF) + ((a >> 16) & 0xFF) + (a >> 24);
}

// Box and threshold of every window of `window` sorted candidates.
static __global__ void k_node_boxes(const uint2 *__restrict__ feat_s, uint32_t N, uint32_t window, uint32_t K,
                                    uint4 *__restrict__ box, uint32_t *__restrict__ thr)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t first = w * window;
    if (first >= N) return;
    const uint32_t last = min(first + window, N);
    uint4 b = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u);
    for (uint32_t k = first; k < last; ++k) {
        const uint2 f = feat_s[k];
        b.x = bytes_min(b.x, f.x);
        b.y = bytes_min(b.y, f.y);
        b.z = bytes_max(b.z, f.x);
        b.w = bytes_max(b.w, f.y);
    }
    box[w] = b;
    thr[w] = bytes_sum(b.z) - bytes_sum(b.x) + bytes_sum(b.w) - bytes_sum(b.y) + 4 * K;
}

} // namespace igc
