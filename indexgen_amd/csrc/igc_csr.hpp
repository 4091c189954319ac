// igc_csr.hpp -- CSR assembly of the conflict graph by an LSD radix sort of
// its directed entries (gfx950).
//
// The pair kernels leave E undirected edges (i << 32) | j, i < j, in no order.
// The CSR wants, for every row r, its neighbours ascending.  Each edge gives two
// directed entries, keyed src << b | dst with b = bits of N-1, so sorting the
// 2E keys orders them by row and then by neighbour.  The sort is LSD radix on
// 8-bit digits, ceil(2b/8) passes, each pass three kernels:
//   k_radix_count    digit histogram of each 4096-entry tile (LDS counters),
//   k_scan_*         exclusive scan of the digit-major counts (the u64 scan the
//                    degree array used to take),
//   k_radix_scatter  stable scatter: within a tile, entries are ranked in input
//                    order by a wave-level multisplit (8 ballots give each lane
//                    the set of lanes holding its digit) plus per-wave digit
//                    counts in LDS.
// Pass 0 reads the edge list directly (entry t is edge t/2, direction t%2).  The
// last pass writes the neighbour (col) and the row (src) of each entry at its
// final place instead of the key, and row_ptr[r] is the first position whose
// src is >= r (binary search).  Every pass moves about 24 B per entry through
// HBM (read for the count, read + write for the scatter) and needs no global
// atomics; equal keys cannot occur (each edge is emitted once).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace igc {

constexpr int kRadixThreads = 256;
constexpr int kRadixRounds = 16; // entries per thread and tile
constexpr uint32_t kRadixTile = kRadixThreads * kRadixRounds;
constexpr int kRadixBits = 8;
constexpr uint32_t kRadixDigits = 1u << kRadixBits;

struct RadixPass {
    const uint64_t *keys;          // input keys (pass > 0)
    const uint64_t *edges;         // edge list (pass 0)
    const unsigned long long *cnt; // edge count on the device
    uint64_t cap;                  // edge buffer capacity (entries beyond are ignored)
    uint32_t bsh;                  // bits of a node id in the key
    uint32_t shift;                // digit position of this pass
    uint32_t from_edges;           // pass 0
};

__device__ __forceinline__ uint64_t radix_items(const RadixPass &p)
{
    const uint64_t e = (uint64_t)*p.cnt;
    return 2 * (e < p.cap ? e : p.cap);
}

__device__ __forceinline__ uint64_t radix_key(const RadixPass &p, uint64_t t)
{
    if (p.from_edges) {
        const uint64_t v = p.edges[t >> 1];
        const uint64_t i = v >> 32, j = (uint32_t)v;
        return (t & 1) ? ((j << p.bsh) | i) : ((i << p.bsh) | j);
    }
    return p.keys[t];
}

// counts[d * tiles + b] = entries of tile b with digit d
static __global__ __launch_bounds__(kRadixThreads) void k_radix_count(RadixPass p, uint32_t tiles,
                                                                       uint32_t *__restrict__ counts)
{
    __shared__ uint32_t h[kRadixDigits];
    const uint32_t b = blockIdx.x;
    for (uint32_t d = threadIdx.x; d < kRadixDigits; d += blockDim.x) h[d] = 0;
    __syncthreads();
    const uint64_t n = radix_items(p);
    const uint64_t start = (uint64_t)b * kRadixTile;
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint64_t t = start + (uint64_t)r * kRadixThreads + threadIdx.x;
        if (t < n) atomicAdd(&h[(uint32_t)(radix_key(p, t) >> p.shift) & (kRadixDigits - 1)], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kRadixDigits; d += blockDim.x) counts[(uint64_t)d * tiles + b] = h[d];
}

// Stable scatter of tile b.  offsets[d * tiles + b] is where the tile's first
// entry with digit d goes.  last: write col/src instead of keys.
static __global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(RadixPass p, uint32_t tiles,
                                                                         const uint64_t *__restrict__ offsets,
                                                                         uint32_t last, uint64_t *__restrict__ out,
                                                                         uint32_t *__restrict__ col,
                                                                         uint32_t *__restrict__ src)
{
    __shared__ uint64_t base[kRadixDigits];
    __shared__ uint32_t wcnt[kRadixThreads / 64][kRadixDigits];
    const uint32_t b = blockIdx.x;
    const uint64_t n = radix_items(p);
    const uint64_t start = (uint64_t)b * kRadixTile;
    if (start >= n) return; // whole workgroup: no barrier is skipped by part of it
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t d = threadIdx.x; d < kRadixDigits; d += blockDim.x) {
        base[d] = offsets[(uint64_t)d * tiles + b];
        for (int w = 0; w < kRadixThreads / 64; ++w) wcnt[w][d] = 0;
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull; // lanes below this one
    const uint64_t mask = (1ull << p.bsh) - 1ull;
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint64_t t = start + (uint64_t)r * kRadixThreads + threadIdx.x;
        const bool valid = t < n;
        const uint64_t key = valid ? radix_key(p, t) : 0ull;
        const uint32_t digit = (uint32_t)(key >> p.shift) & (kRadixDigits - 1);
        // lanes of this wave holding the same digit
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < kRadixBits; ++bit) {
            const bool set = (digit >> bit) & 1u;
            const uint64_t bb = __ballot(set);
            match &= set ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(match & lt);
        if (valid && rank == 0) wcnt[wave][digit] = (uint32_t)__popcll(match);
        __syncthreads();
        if (valid) {
            uint64_t off = base[digit] + rank;
            for (uint32_t w = 0; w < wave; ++w) off += wcnt[w][digit];
            if (last) {
                col[off] = (uint32_t)(key & mask);
                src[off] = (uint32_t)(key >> p.bsh);
            } else {
                out[off] = key;
            }
        }
        __syncthreads();
        for (uint32_t d = threadIdx.x; d < kRadixDigits; d += blockDim.x) {
            uint32_t s = 0;
            for (int w = 0; w < kRadixThreads / 64; ++w) {
                s += wcnt[w][d];
                wcnt[w][d] = 0;
            }
            base[d] += s;
        }
        __syncthreads();
    }
}

// row_ptr[r] = first position with src >= r, for r = 0..N (src ascending).
static __global__ void k_csr_row_ptr(const uint32_t *__restrict__ src, const unsigned long long *__restrict__ cnt,
                                     uint64_t cap, uint32_t N, uint64_t *__restrict__ row_ptr)
{
    const uint64_t e = (uint64_t)*cnt;
    const uint64_t n = 2 * (e < cap ? e : cap);
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= N; r += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (src[mid] < r) lo = mid + 1;
            else hi = mid;
        }
        row_ptr[r] = lo;
    }
}

} // namespace igc
