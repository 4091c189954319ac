// igc_kernels.hpp -- gfx950 kernels of the conflict-graph engine.
//
// Data layout in HBM
//   packed[i]  one word per candidate, two symbol bit planes (igc_lv.hpp):
//              32-bit words for n <= 16 (bit p = sym[p] & 1, bit 16+p =
//              sym[p] >> 1), 64-bit words for 17 <= n <= 32 (planes at bits 0
//              and 32); all other bits 0, so a shift by d moves whole symbols.
//              Binary strings of length 33 <= n <= 64 (symbols 0/1): one
//              64-bit word, bit p = sym[p].
//   edges[e]   (i << 32) | j for every conflicting pair i < j, in no order.
//   row_ptr    uint64 [N+1];  col  uint32 [2|E|]  -- the symmetric CSR.
//
// The pair test.  Edge (i,j) iff Levenshtein(A,B) <= K where K = editDist-1
// (the reference predicate EditDistanceBandedAtLeast, include/EditDistance.hpp
// :368-372, equals min(Levenshtein, K+1) for equal lengths).  Instead of the
// reference's per-column Myers/Hyyro recurrence we run Landau-Vishkin
// furthest-reaching diagonals over mismatch bitmasks:
//   mism_d bit j = (A[j+d] != B[j])   for the 2*floor(K/2)+1 diagonals |d|<=K/2
//   FR[e][d] = slide_d(max(FR[e-1][d]+1, FR[e-1][d-1], FR[e-1][d+1]+1))
//   slide_d(L) = L + ctz(mism_d >> L)
// restricted to |d| <= min(e, K-e), and edge iff FR[K][0] >= n.  Positions
// are held as thermometer codes (mask of bits >= q), so each cell is one AND
// chain, a subtract and one bitwise merge, and for n <= 16 two pairs share a
// register as 16-bit lanes: about 22 VALU instructions per pair at K=3,
// independent of n, instead of the 17*n of one Myers column step per symbol.
// Where that quadratic cost exceeds the Myers recurrence (large K), the
// kernels run Myers instead (pred_kind and the derivation in igc_lv.hpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "igc_lv.hpp"
#include "igc_prefilter.hpp"

namespace igc {

constexpr int kWaves = 4;          // waves per workgroup of the pair kernel
// Edges staged per wave in LDS before a flush.  Every flush takes its slots
// from the one global edge counter, whose atomics serialise on dense graphs
// (4% of pairs conflicting: 256-edge flushes held the pair kernel to a third of
// its issue rate); 512 halves them, and at 16 KB per workgroup eight
// workgroups still fit a CU's LDS.
constexpr int kStage = 512;

// ---------------------------------------------------------------------------
// symbol packing
template <class W>
__global__ void k_pack(const uint8_t *__restrict__ syms, uint32_t N, uint32_t n, W *__restrict__ packed,
                       uint32_t *__restrict__ bad)
{
    constexpr int H = WordTraits<W>::half;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint8_t *s = syms + (size_t)i * n;
    W w = 0;
    uint32_t or_all = 0;
    for (uint32_t p = 0; p < n; ++p) {
        const uint32_t c = s[p];
        or_all |= c;
        w |= (W)(c & 1u) << p;
        w |= (W)((c >> 1) & 1u) << (H + p);
    }
    packed[i] = w;
    if (or_all > 3u) atomicOr(bad, 1u);
}

// Binary strings: one bit per symbol; sets bit 1 of *bad when a symbol exceeds 1.
static __global__ void k_pack_bin(const uint8_t *__restrict__ syms, uint32_t N, uint32_t n,
                                  uint64_t *__restrict__ packed, uint32_t *__restrict__ bad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint8_t *s = syms + (size_t)i * n;
    uint64_t w = 0;
    uint32_t or_all = 0;
    for (uint32_t p = 0; p < n; ++p) {
        const uint32_t c = s[p];
        or_all |= c;
        w |= (uint64_t)(c & 1u) << p;
    }
    packed[i] = w;
    if (or_all > 1u) atomicOr(bad, 2u);
}

// Packed words of uniformly random strings of length n (SplitMix64 per index),
// used to time tile shapes.  BIN: binary words (n <= 64, one plane).
template <class W, bool BIN = false>
__global__ void k_random_packed(W *__restrict__ packed, uint32_t N, uint32_t n, uint64_t seed)
{
    constexpr int H = WordTraits<W>::half;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    if constexpr (BIN) {
        packed[i] = (W)(n >= 64 ? z : (z & ((1ull << n) - 1ull)));
    } else {
        const uint32_t mask = n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
        const uint32_t lo = (uint32_t)z & mask, hi = (uint32_t)(z >> 32) & mask;
        packed[i] = (W)lo | ((W)hi << H);
    }
}

static __global__ void k_fill_u32(uint32_t *__restrict__ p, uint64_t count, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// pair predicates as policies of the tile code.  A policy works on packed words
// W and tests P columns per register (P = 2 on 16-bit lanes for n <= 16, 1 on
// 32-bit lanes above): Col holds a lane's P columns, Row what a wave derives
// once from the (uniform) row word, test() returns the edge lanes as a mask.

template <class W> struct LaneOf;
template <> struct LaneOf<uint32_t> { using type = u16x2; };
template <> struct LaneOf<uint64_t> { using type = uint32_t; };

template <class W>
struct ColPlanes {
    using V = typename LaneOf<W>::type;
    static constexpr int P = LaneCount<V>::value;
    V lo, hi;
    // columns j0 + 64h for h < P (zero words past N)
    __device__ __forceinline__ void load(const W *__restrict__ packed, uint32_t j0, uint32_t N)
    {
        if constexpr (P == 2) {
            const uint32_t w0 = j0 < N ? packed[j0] : 0u;
            const uint32_t w1 = j0 + 64u < N ? packed[j0 + 64u] : 0u;
            col_planes(w0, w1, lo, hi);
        } else {
            col_planes(j0 < N ? packed[j0] : (W)0, lo, hi);
        }
    }
};

template <int K, class W>
struct PolLV {
    using Word = W;
    using V = typename LaneOf<W>::type;
    using Col = ColPlanes<W>;
    static constexpr int P = Col::P;
    using Row = LvtRow<K, V>;
    static __device__ __forceinline__ void row(Row &r, W a, uint32_t, int) { lvt_row<K>(a, r); }
    static __device__ __forceinline__ uint32_t test(const Row &r, const Col &c, uint32_t, int)
    {
        return lvt_test<K, V>(r, c.lo, c.hi);
    }
};

// Landau-Vishkin on 32-bit words (n <= 16, two pairs per register) in row
// coordinates (igc_lv.hpp): the row's planes are the same for every diagonal,
// two scalar packs per row, and each lane keeps its two columns shifted for
// every diagonal, 2(2*floor(K/2)+1) registers loaded once per tile.
template <int K>
struct PolLVR {
    using Word = uint32_t;
    using V = u16x2;
    static constexpr int P = 2;
    struct Col {
        LvtCol<K, u16x2> c;
        __device__ __forceinline__ void load(const uint32_t *__restrict__ packed, uint32_t j0, uint32_t N)
        {
            ColPlanes<uint32_t> b;
            b.load(packed, j0, N);
            lvt_col<K>(b.lo, b.hi, c);
        }
    };
    struct Row {
        u16x2 lo, hi;
    };
    static __device__ __forceinline__ void row(Row &r, uint32_t a, uint32_t, int) { lvt_row_rc(a, r.lo, r.hi); }
    static __device__ __forceinline__ uint32_t test(const Row &r, const Col &c, uint32_t, int)
    {
        return lvt_test_rc<K, u16x2>(r.lo, r.hi, c.c);
    }
};

template <class W>
struct PolMyers {
    using Word = W;
    using V = typename LaneOf<W>::type;
    using Col = ColPlanes<W>;
    static constexpr int P = Col::P;
    struct Row {
        uint32_t lo, hi;
    };
    static __device__ __forceinline__ void row(Row &r, W a, uint32_t, int) { text_planes(a, r.lo, r.hi); }
    static __device__ __forceinline__ uint32_t test(const Row &r, const Col &c, uint32_t n, int k)
    {
        return myers_lanes<V>(r.lo, r.hi, c.lo, c.hi, n, k);
    }
};

// editDist > n: every pair conflicts
template <class W>
struct PolAll {
    using Word = W;
    using Col = ColPlanes<W>;
    static constexpr int P = Col::P;
    struct Row {};
    static __device__ __forceinline__ void row(Row &, W, uint32_t, int) {}
    static __device__ __forceinline__ uint32_t test(const Row &, const Col &, uint32_t, int) { return (1u << P) - 1u; }
};

// Binary words: one column per lane, the word itself.
struct ColBin {
    static constexpr int P = 1;
    uint64_t w;
    __device__ __forceinline__ void load(const uint64_t *__restrict__ packed, uint32_t j0, uint32_t N)
    {
        w = j0 < N ? packed[j0] : 0ull;
    }
};

template <int K>
struct PolLVBin {
    using Word = uint64_t;
    using Col = ColBin;
    static constexpr int P = 1;
    using Row = LvtRowBin<K>;
    static __device__ __forceinline__ void row(Row &r, uint64_t a, uint32_t, int) { lvt_row_bin<K>(a, r); }
    static __device__ __forceinline__ uint32_t test(const Row &r, const Col &c, uint32_t, int)
    {
        return lvt_test_bin<K>(r, c.w);
    }
};

struct PolMyersBin {
    using Word = uint64_t;
    using Col = ColBin;
    static constexpr int P = 1;
    struct Row {
        uint64_t t;
    };
    static __device__ __forceinline__ void row(Row &r, uint64_t a, uint32_t, int) { r.t = a; }
    static __device__ __forceinline__ uint32_t test(const Row &r, const Col &c, uint32_t n, int k)
    {
        return myers_bin(r.t, c.w, n, k);
    }
};

struct PolAllBin {
    using Word = uint64_t;
    using Col = ColBin;
    static constexpr int P = 1;
    struct Row {};
    static __device__ __forceinline__ void row(Row &, uint64_t, uint32_t, int) {}
    static __device__ __forceinline__ uint32_t test(const Row &, const Col &, uint32_t, int) { return 1u; }
};

// The Landau-Vishkin policy of a word class.  32-bit words: row coordinates up
// to K = kLvRowCoordMaxK, where they take the scalar unit off the critical path
// (measured on MI355X, the same session: 3.1 against 3.8 ms for a batch's K = 2
// sets, 5.4 against 5.7 ms for its K = 3 sets), column coordinates above, where
// the kernels are bound by vector issue alone and column coordinates measured
// 1% faster in builds and 3-6% faster count-only.  64-bit two-plane words (one
// pair per 32-bit lane, up to 26 column registers at K = 13 in row coordinates)
// and binary words: column coordinates.
constexpr int kLvRowCoordMaxK = 3;
template <int K, class W, bool BIN> struct LvPolOf {
    using type = PolLV<K, W>;
};
template <int K> struct LvPolOf<K, uint32_t, false> {
    using type = typename std::conditional<(K <= kLvRowCoordMaxK), PolLVR<K>, PolLV<K, uint32_t>>::type;
};
template <int K> struct LvPolOf<K, uint64_t, true> {
    using type = PolLVBin<K>;
};

// ---------------------------------------------------------------------------
// all-pairs tiling of the upper triangle

// Row blocks of R rows, column blocks of W = q*R columns.  Group g holds the q
// row blocks g*q .. g*q+q-1, each owning column blocks g .. NBc-1.
struct TileGeom {
    uint32_t N;
    uint32_t R;
    uint32_t W;
    uint32_t q;
    uint32_t NBr;
    uint32_t NBc;
};

__device__ __forceinline__ uint64_t tiles_before_group(uint64_t g, const TileGeom &G)
{
    return (uint64_t)G.q * (g * (uint64_t)G.NBc - (g * (g - (g > 0 ? 1 : 0))) / 2);
}

__device__ __forceinline__ void decode_tile(uint64_t t, const TileGeom &G, uint32_t &rb, uint32_t &cb)
{
    const double nb = (double)G.NBc + 0.5;
    const double disc = nb * nb - 2.0 * (double)t / (double)G.q;
    int64_t g = (int64_t)(nb - sqrt(disc > 0.0 ? disc : 0.0));
    if (g < 0) g = 0;
    if (g > (int64_t)G.NBc - 1) g = (int64_t)G.NBc - 1;
    while (g > 0 && tiles_before_group((uint64_t)g, G) > t) --g;
    while (g + 1 < (int64_t)G.NBc && tiles_before_group((uint64_t)g + 1, G) <= t) ++g;
    const uint64_t tp = t - tiles_before_group((uint64_t)g, G);
    const uint32_t per = G.NBc - (uint32_t)g;
    // the double-precision estimate runs on the vector unit; the result is
    // uniform, so move it to scalar registers (the row loop bounds follow it)
    rb = __builtin_amdgcn_readfirstlane((uint32_t)g * G.q + (uint32_t)(tp / per));
    cb = __builtin_amdgcn_readfirstlane((uint32_t)g + (uint32_t)(tp % per));
}

__device__ __forceinline__ uint32_t lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t rank_in_mask(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Moves a wave's staged edges to the global edge list with one atomic.
__device__ __forceinline__ void flush_stage(const uint64_t *st, uint32_t staged, uint32_t lane,
                                            uint64_t *__restrict__ edges, uint64_t cap,
                                            unsigned long long *__restrict__ counter)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(counter, (unsigned long long)staged);
    base = uniform64(base);
    for (uint32_t q = lane; q < staged; q += 64) {
        const uint64_t idx = base + q;
        if (idx < cap) edges[idx] = st[q];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Lowest set bit of a wave-uniform mask, -1 when it is empty: s_ff1_i32_b64
// itself (the builtins that define the empty case add a compare and a select).
__device__ __forceinline__ int32_t first_set(uint64_t m)
{
    int32_t t;
    __asm__("s_ff1_i32_b64 %0, %1" : "=s"(t) : "s"(m));
    return t;
}

// m with bit t (0..63) cleared: s_bitset0_b64, one scalar instruction where
// the shift and and-not take two.
__device__ __forceinline__ uint64_t clear_bit(uint64_t m, int32_t t)
{
    __asm__("s_bitset0_b64 %0, %1" : "+s"(m) : "s"(t));
    return m;
}

// Row of a chunk to load for first_set's result t (row 0 when none is left:
// it exists, and the load is not used); the mask (a no-op on 0..63) shows the
// compiler the range, so the scalar load takes its offset in a 32-bit register.
__device__ __forceinline__ uint32_t row_slot(int32_t t) { return (uint32_t)max(t, 0) & 63u; }

// One wave's share of tile (rb, cb) of a candidate set: rows rb*R .. rb*R+R-1,
// columns colw .. colw+64*CPL*P-1 with colw = cb*W + wave*64*CPL*P (lane l,
// slot s -> column colw + 64s + l; slots c*P .. c*P+P-1 share register group c).
// Rows are walked in order with the row word uniform (scalar registers), so
// each lane tests CPL*P pairs per row.  Nonzero ballots (rare) are compacted
// through the wave's LDS stage `st`; edges are written as global ids
// (node_off + i, node_off + j).  EMIT=false only counts.  Contains no
// workgroup barrier: a wave may return early.  n and k are the string length
// and threshold for the policies that take them at run time.
//
// With a RowFilter (single sets), `packed` is in the filter's sorted order, a
// row is skipped for this wave's whole column window when its half-count
// distance to every item of the window exceeds 2K (igc_prefilter.hpp), and
// emitted ids are mapped back to the caller's order through the permutation.
template <class Pol, int CPL, bool EMIT>
__device__ __forceinline__ void wave_tile(const typename Pol::Word *__restrict__ packed, const TileGeom &G,
                                          uint32_t rb, uint32_t cb, uint32_t node_off, uint32_t n, int k,
                                          const RowFilter &flt, uint64_t *st, uint32_t lane, uint32_t wave,
                                          uint64_t *__restrict__ edges, uint64_t cap,
                                          unsigned long long *__restrict__ counter)
{
    constexpr int P = Pol::P;
    constexpr int SL = CPL * P; // columns per lane
    const uint32_t N = G.N;
    const uint32_t row0 = rb * G.R;
    const uint32_t colw = cb * G.W + wave * 64u * SL;
    if (row0 >= N || colw >= N) return;
    const uint32_t colend = min(colw + 64u * SL, N); // exclusive
    // rows i with some column j > i in this wave: i < colend - 1
    if (colend - 1u <= row0) return;
    // uniform; kept in a scalar register so the row loop's exit test is scalar
    const uint32_t rows = __builtin_amdgcn_readfirstlane(min(min(G.R, N - row0), colend - 1u - row0));

    typename Pol::Col col[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) col[c].load(packed, colw + 64u * P * c + lane, N);
    // caller ids of this lane's columns (fixed for the tile), so emitting an
    // edge needs no load
    uint32_t ocol[SL];
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) {
        const uint32_t j = colw + 64u * sl + lane;
        ocol[sl] = (EMIT && flt.perm && j < N) ? flt.perm[j] : j;
    }
    uint32_t staged = 0;
    uint64_t counted = 0;
    const uint4 *items = nullptr;
    const uint32_t *ibias = nullptr;
    uint32_t nitems = 0;
    if (flt.items) {
        const uint32_t w = colw / flt.window;
        items = flt.items + (size_t)w * kItemCap;
        ibias = flt.bias + (size_t)w * kItemCap;
        nitems = __builtin_amdgcn_readfirstlane(flt.count[w]);
    }

    // rows in chunks of 64: with a filter, lane l tests row r0 + l against the
    // window's items and one ballot gives the chunk's rows that can conflict.
    // Each chunk's counts are loaded one chunk ahead, into registers that stay
    // live through the row loop (a load still in flight there would make the
    // row loop wait for every outstanding store as well).
    // With a permutation, the caller ids of a chunk's rows (one per lane) are
    // loaded a chunk ahead too; an emitting row reads its id from them.
    uint2 fnext = make_uint2(0u, 0u);
    uint32_t pnext = 0u;
    if (flt.items && lane < rows) fnext = flt.feat[row0 + lane];
    if (EMIT && flt.perm && lane < rows) pnext = flt.perm[row0 + lane];
    for (uint32_t r0 = 0; r0 < rows; r0 += 64) {
        const uint32_t left = rows - r0;
        uint64_t todo = left >= 64 ? ~0ull : ((1ull << left) - 1ull);
        const uint32_t pids = pnext;
        if (EMIT && flt.perm) {
            const uint32_t nl = r0 + 64u + lane;
            pnext = nl < rows ? flt.perm[row0 + nl] : 0u;
        }
        if (flt.items) {
            const uint2 f = fnext;
            uint32_t best = 0xFFFFFFFFu;
            // the window's items are wave-uniform and read-only: scalar loads
            using CU32 = const __attribute__((address_space(4))) uint32_t;
            CU32 *si = (CU32 *)items, *sb = (CU32 *)ibias;
            // two items a step (slots q0 and q0 + 1, adjacent: kItemCap is even),
            // one min3; the slot is masked to the window's kItemCap (a no-op on
            // q < nitems) so the loads take 32-bit scalar offsets, and a missing
            // second item gets a bias no row can reach
            for (uint32_t q = 0; q < nitems; q += 2) {
                const uint32_t q0 = q & (kItemCap - 2u);
                const uint32_t braw = sb[q0 + 1u]; // in the window's slots either way
                const uint32_t b1 = q + 1u < nitems ? braw : 0x7FFFFFFFu;
                uint32_t v0 = __builtin_amdgcn_sad_u8(f.x, si[4 * q0], sb[q0]);
                uint32_t v1 = __builtin_amdgcn_sad_u8(f.x, si[4 * q0 + 4], b1);
                v0 = __builtin_amdgcn_sad_u8(f.x, si[4 * q0 + 2], v0);
                v1 = __builtin_amdgcn_sad_u8(f.x, si[4 * q0 + 6], v1);
                v0 = __builtin_amdgcn_sad_u8(f.y, si[4 * q0 + 1], v0);
                v1 = __builtin_amdgcn_sad_u8(f.y, si[4 * q0 + 5], v1);
                v0 = __builtin_amdgcn_sad_u8(f.y, si[4 * q0 + 3], v0);
                v1 = __builtin_amdgcn_sad_u8(f.y, si[4 * q0 + 7], v1);
                best = min(best, min(v0, v1));
            }
            todo &= __ballot(best <= flt.limit);
            const uint32_t nl = r0 + 64u + lane;
            fnext = nl < rows ? flt.feat[row0 + nl] : make_uint2(0u, 0u);
        }
        // the next row's word is loaded one row ahead (scalar load latency);
        // the row index and mask stay in scalar registers: t is the next row's
        // offset in the chunk, -1 once none is left (s_ff1's result for an
        // empty mask; the loop tests its sign, and the look-ahead load then
        // reads row r0, which exists, and is not used)
        using Word = typename Pol::Word;
        const Word *rowp = packed + row0 + r0;
        int32_t t = first_set(todo);
        Word next = rowp[row_slot(t)];
        while (t >= 0) {
            const uint32_t i = row0 + r0 + (uint32_t)t;
            todo = clear_bit(todo, t);
            const Word a = next;
            t = first_set(todo);
            next = rowp[row_slot(t)];
            typename Pol::Row row;
            Pol::row(row, a, n, k);
            uint64_t mk[SL];
            uint64_t any = 0;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const uint32_t e = Pol::test(row, col[c], n, k);
#pragma unroll
                for (int h = 0; h < P; ++h) {
                    mk[c * P + h] = __ballot((e >> h) & 1u);
                    any |= mk[c * P + h];
                }
            }
            if (any) {
                // caller id of the row, from the chunk's ids
                const uint32_t oi =
                    (EMIT && flt.perm) ? (uint32_t)__builtin_amdgcn_readlane((int)pids, (int)(i - row0 - r0)) : i;
#pragma unroll
                for (int sl = 0; sl < SL; ++sl) {
                    if (!mk[sl]) continue;
                    const uint32_t j = colw + 64u * sl + lane;
                    const bool mine = ((mk[sl] >> lane) & 1ull) && j > i && j < N;
                    const uint64_t m2 = __ballot(mine);
                    if (!m2) continue;
                    const uint32_t cnt = (uint32_t)__popcll(m2);
                    if constexpr (EMIT) {
                        if (staged + cnt > (uint32_t)kStage) {
                            flush_stage(st, staged, lane, edges, cap, counter);
                            staged = 0;
                        }
                        if (mine)
                            st[staged + rank_in_mask(m2)] = ((uint64_t)(node_off + oi) << 32) | (node_off + ocol[sl]);
                        staged += cnt;
                    } else {
                        counted += cnt;
                    }
                }
            }
        }
    }
    if constexpr (EMIT) {
        if (staged) flush_stage(st, staged, lane, edges, cap, counter);
    } else {
        if (counted && lane == 0) atomicAdd(counter, (unsigned long long)counted);
    }
}

// Single set: one workgroup per tile of the triangle (blockIdx -> tile).
template <class Pol, int CPL, bool EMIT>
__global__ __launch_bounds__(256) void k_pairs(const typename Pol::Word *__restrict__ packed, TileGeom G,
                                               uint32_t node_off, uint32_t n, int k, RowFilter flt,
                                               uint64_t *__restrict__ edges, uint64_t cap,
                                               unsigned long long *__restrict__ counter)
{
    __shared__ uint64_t stage[kWaves][kStage];
    const uint32_t lane = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t rb, cb;
    decode_tile(blockIdx.x, G, rb, cb);
    wave_tile<Pol, CPL, EMIT>(packed, G, rb, cb, node_off, n, k, flt, stage[wave], lane, wave, edges, cap, counter);
}

// ---------------------------------------------------------------------------
// exclusive u64 scan of a u32 array (the radix sort's digit counts)

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive block-wide scan of one uint64 per thread; returns the block total.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *sh /* [kScanThreads/64 + 1] */,
                                                         uint64_t &excl)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += y;
    }
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    uint64_t wave_off = 0, total = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
        if (w < wave) wave_off += sh[w];
        total += sh[w];
    }
    excl = wave_off + incl - v;
    __syncthreads();
    return total;
}

// The kScanItems counts of a thread (base = its first index): four 16-byte
// loads when all of them lie below N (the arrays are 16-byte aligned and base is
// a multiple of kScanItems), else one load per count, 0 past N.
__device__ __forceinline__ void scan_items_load(const uint32_t *__restrict__ a, uint64_t base, uint64_t N,
                                                uint32_t (&v)[kScanItems])
{
    static_assert(kScanItems % 4 == 0, "whole 16-byte loads");
    if (base + kScanItems <= N) {
        const uint4 *p = reinterpret_cast<const uint4 *>(a + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const uint4 x = p[q];
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) v[k] = base + k < N ? a[base + k] : 0u;
    }
}

static __global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *__restrict__ deg, uint32_t N,
                                                               uint64_t *__restrict__ bsum)
{
    __shared__ uint64_t sh[kScanThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    scan_items_load(deg, base, N, v);
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    uint64_t excl;
    const uint64_t tot = block_exclusive_scan(s, sh, excl);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of the nb block sums in place, kScanItems consecutive sums per
// thread and kScanTile per step.
static __global__ __launch_bounds__(kScanThreads) void k_scan_blocks(uint64_t *__restrict__ bsum, uint32_t nb)
{
    __shared__ uint64_t sh[kScanThreads / 64 + 1];
    uint64_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kScanTile) {
        const uint32_t b = b0 + threadIdx.x * kScanItems;
        uint64_t v[kScanItems];
        uint64_t s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = b + k < nb ? bsum[b + k] : 0ull;
            s += v[k];
        }
        uint64_t excl;
        const uint64_t tot = block_exclusive_scan(s, sh, excl);
        uint64_t run = carry + excl;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (b + k < nb) bsum[b + k] = run;
            run += v[k];
        }
        carry += tot;
    }
}

// offsets[i] = boff[block of i] + the counts before i in its block, and
// offsets[N] = the total; a thread's 16 offsets go out as 16-byte stores when
// all of them lie below N.
static __global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t *__restrict__ deg, uint32_t N,
                                                              const uint64_t *__restrict__ boff,
                                                              uint64_t *__restrict__ row_ptr)
{
    __shared__ uint64_t sh[kScanThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    scan_items_load(deg, base, N, v);
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    uint64_t excl;
    block_exclusive_scan(s, sh, excl);
    uint64_t run = boff[blockIdx.x] + excl;
    if (base + kScanItems <= N) {
        ulonglong2 *p = reinterpret_cast<ulonglong2 *>(row_ptr + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 2; ++q) {
            ulonglong2 w;
            w.x = run;
            run += v[2 * q];
            w.y = run;
            run += v[2 * q + 1];
            p[q] = w;
        }
        if (base + kScanItems == N) row_ptr[N] = run;
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < N) {
                row_ptr[base + k] = run;
                run += v[k];
                if (base + k == (uint64_t)N - 1) row_ptr[N] = run;
            }
        }
    }
}

} // namespace igc
