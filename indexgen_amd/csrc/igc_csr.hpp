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
//
// Counting is done per group of lanes, not per entry: the entries of a tile
// come from neighbouring pair tiles, so most of a wave's 64 entries share a
// handful of digits, and one LDS atomic per entry would serialise on the same
// counter.  The count kernels add each run of equal digits across neighbouring
// lanes once (count_runs: a shuffle and a ballot).  The scatters need every
// lane's rank among all lanes of its digit, for stability: the ballots of
// wave_match give the group's size and the rank, the lowest lane of a group
// adds the size, and the ranks are kept for the placement.  Every kernel
// issues all of a thread's loads before using any of them.
//
// Counts and scatters take their tile from xcd_tile: workgroups are dispatched
// to the 8 XCDs round-robin, and consecutive tiles are given to the same XCD
// in order.  A scatter's runs of one digit from consecutive tiles are adjacent
// in the output, and a count kernel's counts of one digit for consecutive
// tiles are adjacent words, so the partial cache lines meet in that XCD's L2
// instead of being written back from two (the counts, written one word per
// tile, otherwise cost a 32-byte sector each).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace igc {

constexpr int kRadixThreads = 512;
constexpr int kRadixRounds = 8; // entries per thread and tile
constexpr uint32_t kRadixTile = kRadixThreads * kRadixRounds;
constexpr uint32_t kRadixMaxBits = 9; // digit width of a pass: 1..9 bits
constexpr uint32_t kRadixMaxDigits = 1u << kRadixMaxBits;

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

// Lanes of this wave holding the same digit as this lane (one ballot per digit
// bit, all kMaxBits of them: bits above the pass's width are zero in every
// lane and leave the set unchanged); lanes with valid == false belong to no
// lane's set.  The lanes that differ from this one in bit k are ballot_k XOR
// m_k, m_k = this lane's bit k spread over the word (0 or ~0); their OR over
// the bits, complemented, is the set.  Per bit: one sign-extending bit
// extract, one compare (the ballot) and one three-input op per 32-bit half.
template <uint32_t kMaxBits>
__device__ __forceinline__ uint64_t wave_match(uint32_t digit, bool valid)
{
    const uint64_t vb = __ballot(valid);
    uint32_t diff_lo = 0, diff_hi = 0;
#pragma unroll
    for (uint32_t bit = 0; bit < kMaxBits; ++bit) {
        uint32_t m = (uint32_t)((int32_t)(digit << (31 - bit)) >> 31); // 0 or ~0
        // opaque, so the ballot compares m itself (not a second shift of digit)
        __asm__ volatile("" : "+v"(m));
        const uint64_t bb = __builtin_amdgcn_ballot_w64(m != 0u);
        // diff |= m ^ ballot half: v_bitop3 with truth table 0xDE = src1 | (src0 ^ src2)
        diff_lo = __builtin_amdgcn_bitop3_b32(m, diff_lo, (uint32_t)bb, 0xDE);
        diff_hi = __builtin_amdgcn_bitop3_b32(m, diff_hi, (uint32_t)(bb >> 32), 0xDE);
    }
    return vb & ~(((uint64_t)diff_hi << 32) | diff_lo);
}

// Adds this round's entries to the histogram h, one atomic per run of equal
// digits in neighbouring lanes.  The valid lanes are a prefix of the wave.
__device__ __forceinline__ void count_runs(uint32_t *h, uint32_t digit, bool valid)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(digit, 1, 64);
    const bool start = valid && (lane == 0 || prev != digit);
    const uint64_t starts = __ballot(start);
    const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));
    if (start) {
        const uint64_t later = starts & ~((2ull << lane) - 1ull); // starts above this lane (none for lane 63)
        const uint32_t end = later ? (uint32_t)__builtin_ctzll(later) : nvalid;
        atomicAdd(&h[digit], end - lane);
    }
}

// Tile of workgroup b of G: XCD x (= b % 8) takes tiles x*q + min(x, G%8) ..
// in order, q = G/8 (a bijection of 0..G-1).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t G)
{
    const uint32_t x = b & 7u, i = b >> 3, q = G >> 3, rem = G & 7u;
    return x * q + (x < rem ? x : rem) + i;
}

// Exclusive scan over the digits of a tile's per-wave counts wpos[w][d] (one
// digit per thread, threads past kDigits own none).  Afterwards wpos[w][d] is
// where wave w's first entry with digit d goes in the tile's digit-ordered
// stage, and the return value is the stage position of digit d's run (for
// the thread owning d).
template <int kThreads, uint32_t kDigits>
__device__ __forceinline__ uint32_t tile_digit_scan(uint32_t (*wpos)[kDigits], uint32_t *wsum)
{
    constexpr int kWaves = kThreads / 64;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t d = threadIdx.x;
    uint32_t c = 0;
    if (d < kDigits)
        for (int w = 0; w < kWaves; ++w) c += wpos[w][d];
    uint32_t incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t run = incl - c;
    for (uint32_t w = 0; w < wave; ++w) run += wsum[w];
    if (d < kDigits) {
        uint32_t pos = run;
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t cw = wpos[w][d];
            wpos[w][d] = pos;
            pos += cw;
        }
    }
    return run;
}

// A lane's result of the counting step, kept for the placement: rank in its
// digit group (6 bits), the group's size (7 bits), the digit (up to 9 bits).
// The empty asm makes the packed word the only value carried to the placement
// (left alone, the compiler keeps the digit and the size in two registers per
// round, 32 more registers for the 32-bit scatter's 16 rounds).
__device__ __forceinline__ uint32_t rank_word(uint32_t rank, uint32_t cnt, uint32_t digit)
{
    uint32_t w = rank | (cnt << 6) | (digit << 13);
    __asm__ volatile("" : "+v"(w));
    return w;
}

// Places one round of a wave's entries (64 consecutive entries of the tile)
// in the digit-ordered stage, from the counting step's rank word.  Every lane
// reads its digit's position before the group's first lane advances it (one
// wave: LDS operations complete in issue order).
template <typename K>
__device__ __forceinline__ void stage_round(K *stage, uint32_t *mypos, bool valid, uint32_t rw, K key)
{
    const uint32_t rank = rw & 63u, cnt = (rw >> 6) & 127u, digit = rw >> 13;
    const uint32_t b0 = valid ? mypos[digit] : 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (valid) {
        stage[b0 + rank] = key;
        if (rank == 0) mypos[digit] = b0 + cnt;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
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
    constexpr uint32_t kPerWave = kSegTile / (kSegThreads / 64);
    __shared__ uint32_t h[kSegMaxDigits];
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
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
    const uint32_t tile_n = (uint32_t)(left < kSegTile ? left : kSegTile);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t key[kSegRounds];
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
        const uint32_t li = wave * kPerWave + (uint32_t)r * 64 + lane;
        key[r] = li < tile_n ? keys[first + li] : 0u;
    }
    for (uint32_t d = threadIdx.x; d < nd; d += blockDim.x) h[d] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r)
        count_runs(h, (key[r] >> shift) & (nd - 1u), wave * kPerWave + (uint32_t)r * 64 + lane < tile_n);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nd; d += blockDim.x)
        counts[(uint64_t)m.tfirst * nd + (uint64_t)d * m.ntiles + k] = h[d];
}

// Stable scatter of one segment tile of 32-bit keys.  Wave w owns the tile's
// entries w*1024 .. w*1024+1023 (16 rounds of 64), so input order within the
// tile is wave-major; each wave counts its digit groups into LDS, one scan
// over the tile's digits gives every wave its running position per digit, the
// waves place their entries in digit order in LDS with the ranks kept from the
// counting, and each digit run is then written out contiguously.  Not last:
// keys go to their padded slots.  Last: the entry at final position g gets
// col[g] = neighbour and src[g] = row, where the key holds (row low bits) <<
// dbits | (neighbour - base[row]) and the row's high bits are the segment.
static __global__ __launch_bounds__(kSegThreads) void k_seg_scatter(
    const uint32_t *__restrict__ keys, const SegMeta *__restrict__ meta, uint32_t nseg,
    const uint32_t *__restrict__ total, uint32_t shift, uint32_t bits, const uint64_t *__restrict__ offsets,
    uint32_t last, uint32_t *__restrict__ out, uint32_t dbits, uint32_t rowbits, const uint32_t *__restrict__ base,
    uint32_t *__restrict__ col, uint32_t *__restrict__ src)
{
    constexpr int kWaves = kSegThreads / 64;
    constexpr uint32_t kPerWave = kSegTile / kWaves;
    static_assert(kSegMaxDigits <= kSegThreads, "one digit per thread");
    __shared__ uint32_t stage[kSegTile];
    __shared__ uint64_t dst[kSegMaxDigits]; // output index of stage position 0, per digit
    __shared__ uint32_t wpos[kWaves][kSegMaxDigits];
    __shared__ uint32_t wsum[kWaves];
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= total[0]) return; // whole workgroup
    const uint32_t sg = seg_of_tile(meta, nseg, t);
    const SegMeta m = meta[sg];
    const uint32_t k = t - m.tfirst;
    const uint64_t first = m.pstart + (uint64_t)k * kSegTile;
    const uint64_t left = m.size - (uint64_t)k * kSegTile;
    const uint32_t tile_n = (uint32_t)(left < kSegTile ? left : kSegTile);
    const uint32_t nd = 1u << bits;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t my_off =
        threadIdx.x < nd ? offsets[(uint64_t)m.tfirst * nd + (uint64_t)threadIdx.x * m.ntiles + k] : 0ull;
    uint32_t key[kSegRounds];
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
        const uint32_t li = wave * kPerWave + (uint32_t)r * 64 + lane;
        key[r] = li < tile_n ? keys[first + li] : 0u;
    }
    for (uint32_t d = threadIdx.x; d < kSegMaxDigits; d += kSegThreads)
        for (int w = 0; w < kWaves; ++w) wpos[w][d] = 0;
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t rw[kSegRounds];
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
        const bool valid = wave * kPerWave + (uint32_t)r * 64 + lane < tile_n;
        const uint32_t digit = (key[r] >> shift) & (nd - 1u);
        const uint64_t match = wave_match<kSegMaxBits>(digit, valid);
        const uint32_t rank = (uint32_t)__popcll(match & lt), cnt = (uint32_t)__popcll(match);
        rw[r] = rank_word(rank, cnt, digit);
        if (valid && rank == 0) atomicAdd(&wpos[wave][digit], cnt);
    }
    __syncthreads();
    {
        const uint32_t run = tile_digit_scan<kSegThreads, kSegMaxDigits>(wpos, wsum);
        // final index = gbase + (stage position - run); padded slot when not last
        if (threadIdx.x < kSegMaxDigits)
            dst[threadIdx.x] = my_off - run + (last ? 0ull : m.pstart - m.ustart);
    }
    __syncthreads();
    uint32_t *mypos = wpos[wave];
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
        const bool valid = wave * kPerWave + (uint32_t)r * 64 + lane < tile_n;
        stage_round<uint32_t>(stage, mypos, valid, rw[r], key[r]);
    }
    __syncthreads();
    const uint32_t dmask = dbits >= 32 ? 0xFFFFFFFFu : ((1u << dbits) - 1u);
#pragma unroll 4
    for (int r = 0; r < kSegRounds; ++r) {
        const uint32_t li = (uint32_t)r * kSegThreads + threadIdx.x;
        if (li < tile_n) {
            const uint32_t kk = stage[li];
            const uint64_t g = dst[(kk >> shift) & (nd - 1u)] + li;
            if (last) {
                const uint32_t row = (sg << rowbits) | (dbits >= 32 ? 0u : (kk >> dbits));
                col[g] = (kk & dmask) + (base ? base[row] : 0u);
                src[g] = row;
            } else {
                out[g] = kk;
            }
        }
    }
}

// Entry of the tile (0..kRadixTile-1) that lane `lane` of wave `wave` holds
// in round r.  Passes over keys keep input order wave-major (the scatter is
// stable).  A first pass from the edge list may take its tile's entries in any
// order (its input order is arbitrary anyway): there round 2s holds the
// direction i -> j of 64 consecutive edges and round 2s+1 the direction j -> i
// of the same edges, so a round's rows come from one side of the pair tiles
// that emitted them.  Either way the valid entries of a round (local index <
// the tile's size) are a prefix of the wave.
__device__ __forceinline__ uint32_t radix_local(const RadixPass &p, uint32_t wave, int r, uint32_t lane)
{
    constexpr uint32_t kPerWave = kRadixTile / (kRadixThreads / 64);
    if (p.from_edges && !p.upper) return wave * kPerWave + ((uint32_t)r >> 1) * 128 + lane * 2 + ((uint32_t)r & 1);
    return wave * kPerWave + (uint32_t)r * 64 + lane;
}

// This thread's keys for rounds r = 0..kRadixRounds-1 (entry start +
// radix_local), every load issued before any is used.  full == false, for a
// count whose digit lies in the row (the split pass): the neighbour field is
// left zero and the set bases are not read.
__device__ __forceinline__ void radix_load(const RadixPass &p, uint64_t start, uint32_t tile_n, bool full,
                                           uint64_t (&key)[kRadixRounds])
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (!p.from_edges) {
#pragma unroll
        for (int r = 0; r < kRadixRounds; ++r) {
            const uint32_t li = radix_local(p, wave, r, lane);
            key[r] = li < tile_n ? p.keys[start + li] : 0ull;
        }
        return;
    }
    uint64_t v[kRadixRounds];
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint32_t li = radix_local(p, wave, r, lane);
        const uint64_t t = start + li;
        v[r] = li < tile_n ? p.edges[p.upper ? t : t >> 1] : 0ull;
    }
    uint32_t b[kRadixRounds];
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint32_t li = radix_local(p, wave, r, lane);
        b[r] = (full && p.base && li < tile_n) ? p.base[v[r] >> 32] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint64_t t = start + radix_local(p, wave, r, lane);
        const uint64_t i = v[r] >> 32, j = (uint32_t)v[r];
        const bool flip = p.upper ? j < i : (t & 1) != 0; // row = j
        const uint64_t row = flip ? j : i, nbr = flip ? i : j;
        key[r] = (row << p.dbits) | (full ? nbr - b[r] : 0ull);
    }
}

// counts[d * tiles + b] = entries of tile b with digit d, d < 2^bits
static __global__ __launch_bounds__(kRadixThreads) void k_radix_count(RadixPass p, uint32_t tiles,
                                                                       uint32_t *__restrict__ counts)
{
    __shared__ uint32_t h[kRadixMaxDigits];
    const uint32_t b = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t nd = 1u << p.bits;
    const uint64_t n = radix_items(p);
    const uint64_t start = (uint64_t)b * kRadixTile;
    const uint32_t tile_n = start >= n ? 0u : (uint32_t)((n - start) < kRadixTile ? (n - start) : kRadixTile);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t key[kRadixRounds];
    radix_load(p, start, tile_n, !(p.from_edges && p.shift >= p.dbits), key);
    for (uint32_t d = threadIdx.x; d < nd; d += blockDim.x) h[d] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r) count_runs(h, radix_digit(p, key[r]), radix_local(p, wave, r, lane) < tile_n);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nd; d += blockDim.x) counts[(uint64_t)d * tiles + b] = h[d];
}

// Stable scatter of tile b; offsets[d * tiles + b] is where the tile's first
// entry with digit d goes.  Wave w owns the tile's entries w*512 .. w*512+511
// (8 rounds of 64), so input order within the tile is wave-major; counting,
// placement in LDS and the contiguous write-out as in k_seg_scatter.  seg: the
// split pass of the segmented form (the low 32 bits of each key go to the
// padded slot of its segment); else last: write col/src instead of keys.
// kBits: the widest digit the instantiation takes (p.bits <= kBits); its LDS
// tables hold 2^kBits digits, so the 8-bit one fits three workgroups per CU
// (42 KB each) where the 9-bit one fits two (54 KB).
template <uint32_t kBits>
static __global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(RadixPass p, uint32_t tiles,
                                                                         const uint64_t *__restrict__ offsets,
                                                                         uint32_t last, uint64_t *__restrict__ out,
                                                                         uint32_t *__restrict__ col,
                                                                         uint32_t *__restrict__ src,
                                                                         const struct SegMeta *__restrict__ seg,
                                                                         uint32_t *__restrict__ out32)
{
    constexpr int kWaves = kRadixThreads / 64;
    constexpr uint32_t kDigits = 1u << kBits;
    static_assert(kBits <= kRadixMaxBits && kDigits <= kRadixThreads, "one digit per thread");
    __shared__ uint64_t stage[kRadixTile];
    __shared__ uint64_t dst[kDigits];          // output index of stage position 0, per digit
    __shared__ uint32_t wpos[kWaves][kDigits]; // per wave: digit count, then next stage position
    __shared__ uint32_t wsum[kWaves];
    const uint32_t b = xcd_tile(blockIdx.x, gridDim.x);
    const uint64_t n = radix_items(p);
    const uint64_t start = (uint64_t)b * kRadixTile;
    if (start >= n) return; // whole workgroup: no barrier is skipped by part of it
    const uint32_t tile_n = (uint32_t)((n - start) < kRadixTile ? (n - start) : kRadixTile);
    const uint32_t nd = 1u << p.bits;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t my_off = threadIdx.x < nd ? offsets[(uint64_t)threadIdx.x * tiles + b] : 0ull;
    uint64_t slot_shift = 0; // segment slot - final position, for the thread's digit
    if (seg && threadIdx.x < nd) slot_shift = seg[threadIdx.x].pstart - seg[threadIdx.x].ustart;
    uint64_t key[kRadixRounds];
    radix_load(p, start, tile_n, true, key);
    for (uint32_t d = threadIdx.x; d < kDigits; d += kRadixThreads)
        for (int w = 0; w < kWaves; ++w) wpos[w][d] = 0;
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t rw[kRadixRounds];
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r) {
        const bool valid = radix_local(p, wave, r, lane) < tile_n;
        const uint32_t digit = radix_digit(p, key[r]);
        const uint64_t match = wave_match<kBits>(digit, valid);
        const uint32_t rank = (uint32_t)__popcll(match & lt), cnt = (uint32_t)__popcll(match);
        rw[r] = rank_word(rank, cnt, digit);
        if (valid && rank == 0) atomicAdd(&wpos[wave][digit], cnt);
    }
    __syncthreads();
    {
        const uint32_t run = tile_digit_scan<kRadixThreads, kDigits>(wpos, wsum);
        if (threadIdx.x < kDigits) dst[threadIdx.x] = my_off - run + slot_shift;
    }
    __syncthreads();
    uint32_t *mypos = wpos[wave];
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r)
        stage_round<uint64_t>(stage, mypos, radix_local(p, wave, r, lane) < tile_n, rw[r], key[r]);
    __syncthreads();
    const uint64_t mask = (1ull << p.bsh) - 1ull;
#pragma unroll 4
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint32_t li = (uint32_t)r * kRadixThreads + threadIdx.x;
        if (li < tile_n) {
            const uint64_t k = stage[li];
            const uint64_t g = dst[radix_digit(p, k)] + li;
            if (seg) {
                out32[g] = (uint32_t)k; // the segment's padded slot
            } else if (last) {
                col[g] = (uint32_t)(k & mask);
                src[g] = (uint32_t)(k >> p.bsh);
            } else {
                out[g] = k;
            }
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
