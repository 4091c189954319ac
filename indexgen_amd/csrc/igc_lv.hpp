// igc_lv.hpp -- the pair predicates of the conflict-graph kernels, written so
// the same code compiles for gfx950 (inside the HIP kernels) and for the host
// (where the tests check it exhaustively against the oracle).
//
// Packed word of a candidate, two symbol planes:
//   32-bit word (n <= 16): bit p = sym[p] & 1, bit 16+p = sym[p] >> 1
//   64-bit word (n <= 32): bit p = sym[p] & 1, bit 32+p = sym[p] >> 1
// all other bits 0.  A shift by d moves whole symbols in both planes.
//
// lv_edge<K>(row shifts of A, B) decides Levenshtein(A, B) <= K by
// Landau-Vishkin furthest-reaching diagonals over mismatch bitmasks (see the
// header of igc_kernels.hpp); myers_edge() decides it with the bit-parallel
// Myers recurrence (one step per symbol of B) for thresholds where
// Landau-Vishkin is costlier or out of its range.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define IGC_HD __host__ __device__ __attribute__((always_inline))
#define IGC_UNROLL _Pragma("unroll")
#else
#define IGC_HD
#define IGC_UNROLL
#endif

namespace igc {

// K value meaning "every pair is an edge" (editDist > n)
constexpr int kAllPairs = 1000;
// largest K lv_edge is defined for
constexpr int kLvMaxK = 15;
// largest K for which pred_kind() picks Landau-Vishkin at any n <= 32 (the
// kernels are instantiated for 0..kLvLaunchMaxK)
constexpr int kLvLaunchMaxK = 9;

template <class W> struct WordTraits;
template <> struct WordTraits<uint32_t> {
    static constexpr int half = 16;
    static constexpr uint32_t hi_ones = 0xFFFF0000u;
    static constexpr uint32_t lo_ones = 0x0000FFFFu;
};
template <> struct WordTraits<uint64_t> {
    static constexpr int half = 32;
    static constexpr uint64_t hi_ones = 0xFFFFFFFF00000000ull;
    static constexpr uint64_t lo_ones = 0x00000000FFFFFFFFull;
};

IGC_HD inline uint32_t ctz_word(uint32_t x) { return (uint32_t)__builtin_ctz(x); }
IGC_HD inline uint32_t ctz_word(uint64_t x) { return (uint32_t)__builtin_ctzll(x); }

IGC_HD inline uint32_t pack_word(const uint8_t *s, uint32_t n)
{
    uint32_t w = 0;
    for (uint32_t p = 0; p < n; ++p) {
        w |= (uint32_t)(s[p] & 1u) << p;
        w |= (uint32_t)((s[p] >> 1) & 1u) << (16 + p);
    }
    return w;
}

IGC_HD inline uint64_t pack_word64(const uint8_t *s, uint32_t n)
{
    uint64_t w = 0;
    for (uint32_t p = 0; p < n; ++p) {
        w |= (uint64_t)(s[p] & 1u) << p;
        w |= (uint64_t)((s[p] >> 1) & 1u) << (32 + p);
    }
    return w;
}

template <int K>
struct LvShape {
    static constexpr int C = (K > kLvMaxK) ? 0 : K / 2; // band half-width
    static constexpr int WD = 2 * C + 1;               // diagonals used
};

// Row word shifted for each diagonal d = t - C (uniform across the wave).
template <int K, class W>
IGC_HD inline void row_shifts(W a, W (&sa)[LvShape<K>::WD])
{
    constexpr int C = LvShape<K>::C;
IGC_UNROLL
    for (int t = 0; t < LvShape<K>::WD; ++t) {
        const int d = t - C;
        sa[t] = d >= 0 ? (W)(a >> d) : (W)(a << (-d));
    }
}

// Levenshtein(A,B) <= K for packed words of one width (see the file header).
// Mismatch mask of diagonal d: low half = symbol mismatches, high half forced
// to 1 so a slide never leaves the word (furthest-reaching values stay below
// half + K + 1 <= width).
template <int K, class W>
IGC_HD inline bool lv_edge(const W (&sa)[LvShape<K>::WD], W b)
{
    constexpr int H = WordTraits<W>::half;
    constexpr W HI = WordTraits<W>::hi_ones;
    constexpr W LO = WordTraits<W>::lo_ones;
    if constexpr (K == 0) {
        const W x = sa[0] ^ b;
        return ((x | (x >> H)) & LO) == 0;
    } else {
        constexpr int C = LvShape<K>::C;
        constexpr int WD = LvShape<K>::WD;
        W m[WD];
        W q0 = 0;
IGC_UNROLL
        for (int t = 0; t < WD; ++t) {
            const W x = sa[t] ^ b;
            const W h = x >> H;
            m[t] = x | h | HI;
            if (t == C) q0 = (x & LO) | h;
        }
        uint32_t fr[WD], fv[WD]; // FR and FR+1 of the previous level
IGC_UNROLL
        for (int t = 0; t < WD; ++t) {
            fr[t] = 0u;
            fv[t] = 0u;
        }
        fr[C] = ctz_word(m[C]);
        fv[C] = fr[C] + 1u;
IGC_UNROLL
        for (int e = 1; e < K; ++e) {
            const int r = (e < K - e) ? e : (K - e);                  // range at level e
            const int rp = (e - 1 < K - e + 1) ? (e - 1) : (K - e + 1); // range at level e-1
            uint32_t nr[WD], nv[WD];
IGC_UNROLL
            for (int t = 0; t < WD; ++t) {
                nr[t] = 0u;
                nv[t] = 0u;
            }
IGC_UNROLL
            for (int d = -C; d <= C; ++d) {
                if (d < -r || d > r) continue;
                uint32_t L = 0u;
                if (d >= -rp && d <= rp) L = fv[d + C];
                if (d - 1 >= -rp && d - 1 <= rp) L = L > fr[d - 1 + C] ? L : fr[d - 1 + C];
                if (d + 1 >= -rp && d + 1 <= rp) L = L > fv[d + 1 + C] ? L : fv[d + 1 + C];
                const uint32_t s = L + ctz_word((W)(m[d + C] >> L));
                nr[d + C] = s;
                nv[d + C] = s + 1u;
            }
IGC_UNROLL
            for (int t = 0; t < WD; ++t) {
                fr[t] = nr[t];
                fv[t] = nv[t];
            }
        }
        uint32_t L = fv[C];
        if constexpr (K >= 2) {
            L = L > fr[C - 1] ? L : fr[C - 1];
            L = L > fv[C + 1] ? L : fv[C + 1];
        }
        return (q0 >> L) == 0;
    }
}

// Bit-parallel Myers (Hyyro's formulation of Myers 1999) deciding
// Levenshtein(T, P) <= K for a text T and a pattern P of one length n <= 32,
// both packed words of one width.  One step per text symbol; the match mask of
// text symbol T[j] is the set of pattern positions p with P_lo[p] == T_lo[j] and
// P_hi[p] == T_hi[j].  Vertical deltas vp/vn of the last column give
// D[n][n] = n + |vp| - |vn| over bits 0..n-1.  Bits at and above n hold
// padding; carries and shifts only move upward, so they never reach the
// counted bits.  In the kernels T is the row (uniform across a wave), so the
// per-step symbol bits are scalar work.
template <class W>
IGC_HD inline bool myers_edge(W t, W p, uint32_t n, int K)
{
    constexpr int H = WordTraits<W>::half;
    constexpr W LO = WordTraits<W>::lo_ones;
    const uint32_t tlo = (uint32_t)(t & LO), thi = (uint32_t)(t >> H);
    const uint32_t plo = (uint32_t)(p & LO), phi = (uint32_t)(p >> H);
    uint32_t vp = ~0u, vn = 0u;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t s0 = 0u - ((tlo >> j) & 1u);
        const uint32_t s1 = 0u - ((thi >> j) & 1u);
        const uint32_t eq = ~((plo ^ s0) | (phi ^ s1));
        const uint32_t d0 = (((eq & vp) + vp) ^ vp) | eq | vn;
        const uint32_t hp = vn | ~(d0 | vp);
        const uint32_t hn = vp & d0;
        const uint32_t x = (hp << 1) | 1u;
        vn = x & d0;
        vp = (hn << 1) | ~(x | d0);
    }
    const uint32_t mask = n >= 32 ? ~0u : ((1u << n) - 1u);
    const int dist = (int)n + __builtin_popcount(vp & mask) - __builtin_popcount(vn & mask);
    return dist <= K;
}

// Which predicate a (length, threshold) pair uses: Landau-Vishkin on the word
// width of n while its cost (VALU instructions per pair) is below Myers', the
// Myers recurrence otherwise, and no test at all when K >= n.
enum class PredKind : int { LV32 = 0, LV64 = 1, MYERS = 2, ALL = 3 };

// VALU instructions of one pair test on gfx950, counted in the hipcc -O3
// assembly of a kernel that runs just the predicate (v_bitop3_b32 merges most
// of the Myers step's logic): Landau-Vishkin per K on 32- and 64-bit words,
// Myers 15 per symbol plus 10.
IGC_HD inline uint32_t lv_cost(int K, bool wide)
{
    const uint32_t c32[kLvMaxK + 1] = {6, 10, 26, 40, 67, 91, 128, 162, 209, 253, 310, 364, 431, 495, 572, 646};
    const uint32_t c64[kLvMaxK + 1] = {8, 12, 40, 63, 107, 146, 206, 261, 337, 408, 500, 587, 695, 798, 922, 1041};
    return wide ? c64[K] : c32[K];
}

IGC_HD inline uint32_t myers_cost(uint32_t n) { return 15u * n + 10u; }

IGC_HD inline PredKind pred_kind(uint32_t n, int K)
{
    if (K >= kAllPairs || K >= (int)n) return PredKind::ALL;
    const bool wide = n > 16;
    if (K <= kLvMaxK && lv_cost(K, wide) <= myers_cost(n)) return wide ? PredKind::LV64 : PredKind::LV32;
    return PredKind::MYERS;
}

} // namespace igc
