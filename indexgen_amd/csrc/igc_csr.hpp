// igc_csr.hpp -- CSR assembly of the conflict graph by an LSD radix sort of
// its directed entries (gfx950).
//
// The pair kernels leave E undirected edges (i << 32) | j, i < j, in no order.
// The CSR wants, for every row r, its neighbours ascending.  Each edge gives two
// directed entries, keyed src << b | dst with b = bits of N-1, so sorting the
// 2E keys orders them by row and then by neighbour.  The sort is LSD radix in
// P = ceil(2b/9) passes of ceil(2b/P)-bit digits (at most 512 digits), each pass
// three kernels:
//   k_radix_count    digit histogram of each 4096-entry tile (LDS counters),
//   k_scan_*         exclusive scan of the digit-major counts (the u64 scan the
//                    degree array used to take),
//   k_radix_scatter  stable scatter: within a tile, entries are ranked in input
//                    order by a wave-level multisplit (one ballot per digit bit
//                    gives each lane the set of lanes holding its digit) plus per-wave digit
//                    counts in LDS, placed in digit order in LDS, and written
//                    out one contiguous digit run after another.
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

constexpr int kRadixThreads = 512;
constexpr int kRadixRounds = 8; // entries per thread and tile
constexpr uint32_t kRadixTile = kRadixThreads * kRadixRounds;
constexpr uint32_t kRadixMaxBits = 9; // digit width of a pass: 1..9 bits
constexpr uint32_t kRadixMaxDigits = 1u << kRadixMaxBits;
constexpr uint32_t kDigitsPerThread = kRadixMaxDigits / kRadixThreads;

struct RadixPass {
    const uint64_t *keys;          // input keys (pass > 0)
    const uint64_t *edges;         // edge list (pass 0)
    const unsigned long long *cnt; // edge count on the device
    uint64_t cap;                  // edge buffer capacity (entries beyond are ignored)
    uint32_t bsh;                  // bits of a node id (the row) in the key
    uint32_t dbits;                // bits of the neighbour field (below the row)
    const uint32_t *base;          // neighbour field = neighbour - base[row] (null: 0)
    uint32_t shift;                // digit position of this pass
    uint32_t bits;                 // digit width of this pass
    uint32_t from_edges;           // pass 0
    uint32_t upper;                // one entry per edge, row = the smaller id
};

__device__ __forceinline__ uint64_t radix_items(const RadixPass &p)
{
    const uint64_t e = (uint64_t)*p.cnt;
    return (p.upper ? 1 : 2) * (e < p.cap ? e : p.cap);
}

// Key of directed entry t: row << dbits | (neighbour - base), where base is
// the first node of the edge's set (both ends of an edge lie in one set, so
// the lookup is by the edge's first id: both directions of an edge, in
// neighbouring lanes, read the same word).
__device__ __forceinline__ uint64_t radix_key(const RadixPass &p, uint64_t t)
{
    if (p.from_edges) {
        const uint64_t v = p.edges[p.upper ? t : t >> 1];
        const uint64_t i = v >> 32, j = (uint32_t)v;
        const uint64_t b = p.base ? p.base[i] : 0;
        const bool flip = p.upper ? j < i : (t & 1) != 0; // row = j
        return flip ? ((j << p.dbits) | (i - b)) : ((i << p.dbits) | (j - b));
    }
    return p.keys[t];
}

__device__ __forceinline__ uint32_t radix_digit(const RadixPass &p, uint64_t key)
{
    return (uint32_t)(key >> p.shift) & ((1u << p.bits) - 1u);
}

// ---------------------------------------------------------------------------
// Segmented form for large graphs: the 64-bit keys row << dbits | neighbour
// are first split by their top bits (the row's high part; at most 9 bits) in
// one pass straight from the edge list, which writes each entry's low 32 bits
// into its segment; each segment starts at a multiple of kSegTile slots
// (padding slots are never read).  The remaining passes sort 32-bit keys
// segment by segment: a tile never straddles two segments, its digit counts
// are laid out [segment][digit][tile of the segment], so one global scan
// gives every (tile, digit) its final position, and only 4 bytes per entry
// move in each of them.
constexpr int kSegThreads = 512;
constexpr int kSegRounds = 16;
constexpr uint32_t kSegTile = kSegThreads * kSegRounds; // 32-bit keys per tile
constexpr uint32_t kSegMaxBits = 8;                     // digits of the 32-bit passes (4 passes of 8 bits)
constexpr uint32_t kSegMaxDigits = 1u << kSegMaxBits;

struct SegMeta {
    uint64_t ustart; // first final position of the segment
    uint64_t pstart; // first padded slot
    uint64_t size;   // entries
    uint32_t tfirst; // first tile
    uint32_t ntiles;
};

__device__ __forceinline__ uint64_t seg_slot(const SegMeta &m, uint64_t g) { return g - m.ustart + m.pstart; }

// Segment table from the split pass's digit-major offsets (tiles64 tiles,
// nseg digits): one workgroup, one thread per segment; total[0] = tiles of
// all segments.
static __global__ __launch_bounds__(512) void k_seg_meta(const uint64_t *__restrict__ offsets, uint32_t tiles64,
                                                         uint32_t nseg, SegMeta *__restrict__ meta,
                                                         uint32_t *__restrict__ total)
{
    __shared__ uint32_t nt[512];
    const uint32_t d = threadIdx.x;
    uint64_t us = 0, sz = 0;
    if (d < nseg) {
        us = offsets[(uint64_t)d * tiles64];
        sz = offsets[(uint64_t)(d + 1) * tiles64] - us;
    }
    nt[d] = (uint32_t)((sz + kSegTile - 1) / kSegTile);
    __syncthreads();
    if (d == 0) { // exclusive scan of at most 512 tile counts
        uint32_t run = 0;
        for (uint32_t k = 0; k < nseg; ++k) {
            const uint32_t c = nt[k];
            nt[k] = run;
            run += c;
        }
        total[0] = run;
    }
    __syncthreads();
    if (d < nseg) {
        const uint32_t tf = nt[d];
        meta[d] = SegMeta{us, (uint64_t)tf * kSegTile, sz, tf, (uint32_t)((sz + kSegTile - 1) / kSegTile)};
    }
}

// Segment of tile t: the last segment whose first tile is <= t (segments
// without tiles share their successor's first tile and are skipped).
__device__ __forceinline__ uint32_t seg_of_tile(const SegMeta *__restrict__ meta, uint32_t nseg, uint32_t t)
{
    uint32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (meta[mid].tfirst <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// counts[tfirst*nd + d*ntiles + k] = entries with digit d in tile k of the
// segment (32-bit keys, digit = key >> shift, `bits` wide).  Workgroups past
// the last tile zero the rest of the counts (up to gridDim.x * nd), which the
// scan covers.
static __global__ __launch_bounds__(kSegThreads) void k_seg_count(const uint32_t *__restrict__ keys,
                                                                   const SegMeta *__restrict__ meta, uint32_t nseg,
                                                                   const uint32_t *__restrict__ total, uint32_t shift,
                                                                   uint32_t bits, uint32_t *__restrict__ counts)
{
    __shared__ uint32_t h[kSegMaxDigits];
    const uint32_t t = blockIdx.x;
    const uint32_t nd = 1u << bits;
    const uint32_t tt = total[0];
    if (t >= tt) {
        for (uint32_t d = threadIdx.x; d < nd; d += blockDim.x) counts[(uint64_t)t * nd + d] = 0u;
        return; // whole workgroup
    }
    const uint32_t sg = seg_of_tile(meta, nseg, t);
    const SegMeta m = meta[sg];
    const uint32_t k = t - m.tfirst;
    const uint64_t first = m.pstart + (uint64_t)k * kSegTile;
    const uint64_t left = m.size - (uint64_t)k * kSegTile;
    const uint32_t n = (uint32_t)(left < kSegTile ? left : kSegTile);
    for (uint32_t d = threadIdx.x; d < nd; d += blockDim.x) h[d] = 0;
    __syncthreads();
    for (uint32_t li = threadIdx.x; li < n; li += kSegThreads)
        atomicAdd(&h[(keys[first + li] >> shift) & (nd - 1u)], 1u);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nd; d += blockDim.x)
        counts[(uint64_t)m.tfirst * nd + (uint64_t)d * m.ntiles + k] = h[d];
}

// Stable scatter of one segment tile of 32-bit keys (ranking as in
// k_radix_scatter).  Not last: keys go to their padded slots.  Last: the entry
// at final position g gets col[g] = neighbour and src[g] = row, where the key
// holds (row low bits) << dbits | (neighbour - base[row]) and the row's high
// bits are the segment.
static __global__ __launch_bounds__(kSegThreads) void k_seg_scatter(
    const uint32_t *__restrict__ keys, const SegMeta *__restrict__ meta, uint32_t nseg,
    const uint32_t *__restrict__ total, uint32_t shift, uint32_t bits, const uint64_t *__restrict__ offsets,
    uint32_t last, uint32_t *__restrict__ out, uint32_t dbits, uint32_t rowbits, const uint32_t *__restrict__ base,
    uint32_t *__restrict__ col, uint32_t *__restrict__ src)
{
    constexpr int kWaves = kSegThreads / 64;
    constexpr uint32_t kPerWave = kSegTile / kWaves;
    constexpr uint32_t kDigitsPerThreadS = (kSegMaxDigits + kSegThreads - 1) / kSegThreads; // 1
    __shared__ uint32_t stage[kSegTile];
    __shared__ uint64_t gbase[kSegMaxDigits];
    __shared__ uint32_t lstart[kSegMaxDigits];
    __shared__ uint32_t wpos[kWaves][kSegMaxDigits];
    __shared__ uint32_t wsum[kWaves];
    const uint32_t t = blockIdx.x;
    if (t >= total[0]) return; // whole workgroup
    const uint32_t sg = seg_of_tile(meta, nseg, t);
    const SegMeta m = meta[sg];
    const uint32_t k = t - m.tfirst;
    const uint64_t first = m.pstart + (uint64_t)k * kSegTile;
    const uint64_t left = m.size - (uint64_t)k * kSegTile;
    const uint32_t tile_n = (uint32_t)(left < kSegTile ? left : kSegTile);
    const uint32_t nd = 1u << bits;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // this thread's digit offset, loaded now so its latency overlaps the counting
    const uint64_t my_off =
        threadIdx.x < nd ? offsets[(uint64_t)m.tfirst * nd + (uint64_t)threadIdx.x * m.ntiles + k] : 0ull;
    for (uint32_t d = threadIdx.x; d < kSegMaxDigits; d += kSegThreads)
        for (int w = 0; w < kWaves; ++w) wpos[w][d] = 0;
    __syncthreads();
    uint32_t key[kSegRounds];
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
        const uint32_t li = wave * kPerWave + (uint32_t)r * 64 + lane;
        key[r] = li < tile_n ? keys[first + li] : 0u;
        if (li < tile_n) atomicAdd(&wpos[wave][(key[r] >> shift) & (nd - 1u)], 1u);
    }
    __syncthreads();
    {
        // thread d owns digit d (threads past the digits hold zeros)
        uint32_t c[kDigitsPerThreadS], tsum = 0;
#pragma unroll
        for (uint32_t q = 0; q < kDigitsPerThreadS; ++q) {
            const uint32_t d = threadIdx.x * kDigitsPerThreadS + q;
            c[q] = 0;
            if (d < kSegMaxDigits) {
                for (int w = 0; w < kWaves; ++w) c[q] += wpos[w][d];
                gbase[d] = my_off; // d == threadIdx.x (one digit per thread)
            }
            tsum += c[q];
        }
        uint32_t incl = tsum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= (uint32_t)o) incl += y;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint32_t run = incl - tsum;
        for (uint32_t w = 0; w < wave; ++w) run += wsum[w];
#pragma unroll
        for (uint32_t q = 0; q < kDigitsPerThreadS; ++q) {
            const uint32_t d = threadIdx.x * kDigitsPerThreadS + q;
            if (d >= kSegMaxDigits) break;
            lstart[d] = run;
            uint32_t pos = run;
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t cw = wpos[w][d];
                wpos[w][d] = pos;
                pos += cw;
            }
            run += c[q];
        }
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t *mypos = wpos[wave];
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
        const uint32_t li = wave * kPerWave + (uint32_t)r * 64 + lane;
        const bool valid = li < tile_n;
        const uint32_t digit = (key[r] >> shift) & (nd - 1u);
        uint64_t match = __ballot(valid);
#pragma unroll
        for (uint32_t bit = 0; bit < kSegMaxBits; ++bit) {
            if (bit >= bits) break; // uniform
            const bool set = (digit >> bit) & 1u;
            const uint64_t bb = __ballot(set);
            match &= set ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(match & lt);
        const uint32_t b0 = valid ? mypos[digit] : 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            stage[b0 + rank] = key[r];
            if (rank == 0) mypos[digit] = b0 + (uint32_t)__popcll(match);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    const uint32_t dmask = dbits >= 32 ? 0xFFFFFFFFu : ((1u << dbits) - 1u);
    for (uint32_t li = threadIdx.x; li < tile_n; li += kSegThreads) {
        const uint32_t kk = stage[li];
        const uint32_t digit = (kk >> shift) & (nd - 1u);
        const uint64_t g = gbase[digit] + (li - lstart[digit]);
        if (last) {
            const uint32_t row = (sg << rowbits) | (dbits >= 32 ? 0u : (kk >> dbits));
            col[g] = (kk & dmask) + (base ? base[row] : 0u);
            src[g] = row;
        } else {
            out[seg_slot(m, g)] = kk;
        }
    }
}

// counts[d * tiles + b] = entries of tile b with digit d, d < 2^bits
static __global__ __launch_bounds__(kRadixThreads) void k_radix_count(RadixPass p, uint32_t tiles,
                                                                       uint32_t *__restrict__ counts)
{
    __shared__ uint32_t h[kRadixMaxDigits];
    const uint32_t b = blockIdx.x;
    const uint32_t nd = 1u << p.bits;
    for (uint32_t d = threadIdx.x; d < nd; d += blockDim.x) h[d] = 0;
    __syncthreads();
    const uint64_t n = radix_items(p);
    const uint64_t start = (uint64_t)b * kRadixTile;
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint64_t t = start + (uint64_t)r * kRadixThreads + threadIdx.x;
        if (t < n) atomicAdd(&h[radix_digit(p, radix_key(p, t))], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nd; d += blockDim.x) counts[(uint64_t)d * tiles + b] = h[d];
}

// Stable scatter of tile b; offsets[d * tiles + b] is where the tile's first
// entry with digit d goes.  Wave w owns the tile's entries w*512 .. w*512+511
// (8 rounds of 64), so input order within the tile is wave-major.  Each wave
// counts its digits into LDS; one scan over the tile's digits gives every wave
// its running position per digit; the waves then rank their entries round by
// round with wave-level operations only (one ballot per digit bit gives each
// lane the lanes holding its digit), placing the tile in digit order in LDS;
// finally each digit run is written out contiguously by consecutive threads.
// last: write col/src instead of keys.
static __global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(RadixPass p, uint32_t tiles,
                                                                         const uint64_t *__restrict__ offsets,
                                                                         uint32_t last, uint64_t *__restrict__ out,
                                                                         uint32_t *__restrict__ col,
                                                                         uint32_t *__restrict__ src,
                                                                         const struct SegMeta *__restrict__ seg,
                                                                         uint32_t *__restrict__ out32)
{
    constexpr int kWaves = kRadixThreads / 64;
    constexpr uint32_t kPerWave = kRadixTile / kWaves;
    __shared__ uint64_t stage[kRadixTile];
    __shared__ uint64_t gbase[kRadixMaxDigits];        // global position of the tile's first entry per digit
    __shared__ uint32_t lstart[kRadixMaxDigits];       // local position of the same
    __shared__ uint32_t wpos[kWaves][kRadixMaxDigits]; // per wave: digit count, then next local position
    __shared__ uint32_t wsum[kWaves];
    const uint32_t b = blockIdx.x;
    const uint64_t n = radix_items(p);
    const uint64_t start = (uint64_t)b * kRadixTile;
    if (start >= n) return; // whole workgroup: no barrier is skipped by part of it
    const uint32_t tile_n = (uint32_t)((n - start) < kRadixTile ? (n - start) : kRadixTile);
    const uint32_t nd = 1u << p.bits;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    static_assert(kDigitsPerThread == 1, "one digit per thread");
    // this thread's digit offset, loaded now so its latency overlaps the counting
    const uint64_t my_off = threadIdx.x < nd ? offsets[(uint64_t)threadIdx.x * tiles + b] : 0ull;
    for (uint32_t d = threadIdx.x; d < kRadixMaxDigits; d += kRadixThreads)
        for (int w = 0; w < kWaves; ++w) wpos[w][d] = 0;
    __syncthreads();
    uint64_t key[kRadixRounds];
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint32_t li = wave * kPerWave + (uint32_t)r * 64 + lane;
        key[r] = li < tile_n ? radix_key(p, start + li) : 0ull;
        if (li < tile_n) atomicAdd(&wpos[wave][radix_digit(p, key[r])], 1u);
    }
    __syncthreads();
    {
        // tile counts, their exclusive scan over digits, and each wave's start;
        // thread t owns digits t*kDigitsPerThread .. +kDigitsPerThread-1
        uint32_t c[kDigitsPerThread], tsum = 0;
#pragma unroll
        for (uint32_t k = 0; k < kDigitsPerThread; ++k) {
            const uint32_t d = threadIdx.x * kDigitsPerThread + k;
            c[k] = 0;
            for (int w = 0; w < kWaves; ++w) c[k] += wpos[w][d];
            gbase[d] = my_off; // d == threadIdx.x
            tsum += c[k];
        }
        uint32_t incl = tsum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= (uint32_t)o) incl += y;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint32_t run = incl - tsum;
        for (uint32_t w = 0; w < wave; ++w) run += wsum[w];
#pragma unroll
        for (uint32_t k = 0; k < kDigitsPerThread; ++k) {
            const uint32_t d = threadIdx.x * kDigitsPerThread + k;
            lstart[d] = run;
            uint32_t pos = run;
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t cw = wpos[w][d];
                wpos[w][d] = pos;
                pos += cw;
            }
            run += c[k];
        }
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull; // lanes below this one
    uint32_t *mypos = wpos[wave];
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint32_t li = wave * kPerWave + (uint32_t)r * 64 + lane;
        const bool valid = li < tile_n;
        const uint32_t digit = radix_digit(p, key[r]);
        uint64_t match = __ballot(valid);
#pragma unroll
        for (uint32_t bit = 0; bit < kRadixMaxBits; ++bit) {
            if (bit >= p.bits) break; // uniform
            const bool set = (digit >> bit) & 1u;
            const uint64_t bb = __ballot(set);
            match &= set ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(match & lt);
        // every lane reads its digit's position before the group's first lane
        // advances it (one wave: LDS operations complete in issue order)
        const uint32_t base = valid ? mypos[digit] : 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            stage[base + rank] = key[r];
            if (rank == 0) mypos[digit] = base + (uint32_t)__popcll(match);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // global position = gbase[d] + (local position - start of digit d's run)
    const uint64_t mask = (1ull << p.bsh) - 1ull;
    for (uint32_t li = threadIdx.x; li < tile_n; li += kRadixThreads) {
        const uint64_t k = stage[li];
        const uint32_t digit = radix_digit(p, k);
        const uint64_t g = gbase[digit] + (li - lstart[digit]);
        if (seg) {
            out32[seg_slot(seg[digit], g)] = (uint32_t)k; // the segment's padded slot
        } else if (last) {
            col[g] = (uint32_t)(k & mask);
            src[g] = (uint32_t)(k >> p.bsh);
        } else {
            out[g] = k;
        }
    }
}

// row_ptr[r] = first position with src >= r, for r = 0..N (src ascending).
static __global__ void k_csr_row_ptr(const uint32_t *__restrict__ src, const unsigned long long *__restrict__ cnt,
                                     uint64_t cap, uint32_t per_edge, uint32_t N, uint64_t *__restrict__ row_ptr)
{
    const uint64_t e = (uint64_t)*cnt;
    const uint64_t n = per_edge * (e < cap ? e : cap);
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
