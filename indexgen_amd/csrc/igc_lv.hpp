// igc_lv.hpp -- the pair predicates of the conflict-graph kernels, written so
// the same code compiles for gfx950 (inside the HIP kernels) and for the host
// (where the tests check it exhaustively against the oracle).
//
// Packed word of a candidate, two symbol planes:
//   32-bit word (n <= 16): bit p = sym[p] & 1, bit 16+p = sym[p] >> 1
//   64-bit word (n <= 32): bit p = sym[p] & 1, bit 32+p = sym[p] >> 1
// all other bits 0.  A shift by d moves whole symbols in both planes.
// Binary strings (symbols 0/1) of length 33 <= n <= 64 use one plane:
//   64-bit binary word: bit p = sym[p], all other bits 0.
//
// The tests run on a lane type V: u16x2 holds two pairs, one per 16-bit half
// of a 32-bit register (n <= 16; packed 16-bit subtract, add and shift, bitwise
// ops on the whole register), uint32_t holds one pair (n <= 32).  Both tests
// give Levenshtein(A, B) <= K exactly (binary words run the same tests on
// uint64_t lanes, one pair per register pair):
//   lvt_test     Landau-Vishkin furthest-reaching diagonals on thermometer
//                codes, O(K^2) instructions independent of n;
//   myers_lanes  the bit-parallel Myers recurrence, O(n) instructions.
// pred_kind() picks the faster one for (n, K) by measured times.
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
// largest K lvt_residue is defined for
constexpr int kLvMaxK = 15;
// largest K the Landau-Vishkin kernels are instantiated for; pred_kind() picks
// Myers above it
constexpr int kLvLaunchMaxK = 13;

template <class W> struct WordTraits;
template <> struct WordTraits<uint32_t> {
    static constexpr int half = 16;
};
template <> struct WordTraits<uint64_t> {
    static constexpr int half = 32;
};

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

IGC_HD inline uint64_t pack_bin64(const uint8_t *s, uint32_t n)
{
    uint64_t w = 0;
    for (uint32_t p = 0; p < n; ++p) w |= (uint64_t)(s[p] & 1u) << p;
    return w;
}

// ---------------------------------------------------------------------------
// lane types

#if defined(__clang__)
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
#else
typedef uint16_t u16x2 __attribute__((vector_size(4)));
#endif

template <class V> struct LaneCount;
template <> struct LaneCount<uint32_t> { static constexpr int value = 1; };
template <> struct LaneCount<u16x2> { static constexpr int value = 2; };

IGC_HD inline u16x2 dup16(uint32_t x)
{
    u16x2 v;
    v[0] = (uint16_t)x;
    v[1] = (uint16_t)x;
    return v;
}

template <class V> IGC_HD inline V splat_lane(uint32_t x);
template <> IGC_HD inline uint32_t splat_lane<uint32_t>(uint32_t x) { return x; }
template <> IGC_HD inline u16x2 splat_lane<u16x2>(uint32_t x) { return dup16(x); }

IGC_HD inline uint32_t lane_bits(uint32_t v, int) { return v; }
IGC_HD inline uint32_t lane_bits(u16x2 v, int h) { return (uint32_t)v[h]; }

// lanes whose value is zero, as a bit mask
IGC_HD inline uint32_t zero_lanes(uint32_t z) { return z == 0 ? 1u : 0u; }
IGC_HD inline uint32_t zero_lanes(u16x2 z) { return (z[0] == 0 ? 1u : 0u) | (z[1] == 0 ? 2u : 0u); }
IGC_HD inline uint32_t zero_lanes(uint64_t z) { return z == 0 ? 1u : 0u; }

// Planes of one row word shifted by s (s >= 0: towards bit 0), one lane's width.
IGC_HD inline void row_planes(uint32_t a, int s, u16x2 &lo, u16x2 &hi)
{
    const uint32_t l = a & 0xFFFFu, h = a >> 16;
    lo = dup16(s >= 0 ? (l >> s) : ((l << (-s)) & 0xFFFFu));
    hi = dup16(s >= 0 ? (h >> s) : ((h << (-s)) & 0xFFFFu));
}
IGC_HD inline void row_planes(uint64_t a, int s, uint32_t &lo, uint32_t &hi)
{
    const uint32_t l = (uint32_t)a, h = (uint32_t)(a >> 32);
    lo = s >= 0 ? (l >> s) : (l << (-s));
    hi = s >= 0 ? (h >> s) : (h << (-s));
}

// Column planes: two 32-bit packed words, one per 16-bit lane, or one 64-bit word.
IGC_HD inline void col_planes(uint32_t w0, uint32_t w1, u16x2 &lo, u16x2 &hi)
{
    lo[0] = (uint16_t)w0;
    lo[1] = (uint16_t)w1;
    hi[0] = (uint16_t)(w0 >> 16);
    hi[1] = (uint16_t)(w1 >> 16);
}
IGC_HD inline void col_planes(uint64_t w, uint32_t &lo, uint32_t &hi)
{
    lo = (uint32_t)w;
    hi = (uint32_t)(w >> 32);
}

// ---------------------------------------------------------------------------
// Landau-Vishkin on thermometer codes
//
// Diagonal d compares A[j+d] with B[j]; its mismatch mask has bit j set iff
// they differ.  FR[e][d], the furthest position of diagonal d reachable with e
// edits, obeys
//   FR[e][d] = slide_d(max(FR[e-1][d] + 1, FR[e-1][d-1], FR[e-1][d+1] + 1))
// with slide_d(L) = the first mismatch of diagonal d at or after L, restricted
// to |d| <= min(e, K-e) (a path through diagonal d to diagonal 0 needs |d| edits
// each way), and Levenshtein(A, B) <= K iff slide_0 at level K reaches n.
// A position q is held as its thermometer code, the mask of all bits >= q
// (0 = past the lane), so
//   max of positions   = AND of their masks,
//   position + 1       = mask << 1,
//   slide from mask L  = ~(x - 1) | x  with x = m & L (its lowest set bit and
//                        everything above; 0 when no mismatch is left),
// which needs no count-trailing-zeros and no max.  Positions past a diagonal's
// end (garbage bits of the shifted word) only ever mean "end reached": the DP
// cells of neighbouring diagonals differ by at most one.  On diagonal 0 the
// padding positions n.. match (both strings are zero-padded), so the final
// residue m_0 & L_K is zero iff the last slide reaches the lane's end.
//
// Row coordinates (RC).  The same diagonals indexed by the row position
// i = j + d instead: bit i of diagonal d's mask is A[i] != B[i-d], so A's planes
// are the same for every diagonal and B's planes carry the shift (B << d, or
// B >> -d).  A position q of diagonal d moves to q + d, so of the three terms
// only the neighbours' change: FR[e-1][d-1] lands one bit further on diagonal
// d (its mask << 1) and FR[e-1][d+1] + 1 lands on the same bit (its mask as it
// is).  The diagonals' valid ranges shift with them: diagonal d > 0 starts at
// bit d, and every mask it is given is at least FR[e-1][d-1] shifted there
// (d - 1 is in range whenever d is), so its garbage bits below d (B's shifted-in
// zeros) are never read; its end is bit n.  Diagonal d < 0 runs over bits
// 0 .. n+d-1 and, as before, reads nothing below its start.  Every lane width
// holds these ranges for any n the lane takes.  In the kernels the row is
// uniform across a wave and the columns are fixed for a whole tile, so row
// coordinates take every per-row shift off the scalar unit: a lane shifts its
// columns once per tile.

template <int K>
struct LvShape {
    static constexpr int C = (K > kLvMaxK) ? 0 : K / 2; // band half-width
    static constexpr int WD = 2 * C + 1;               // diagonals used
};

// m[t] is the mismatch mask of diagonal d = t - C, in column coordinates, or
// in row coordinates when RC; returns the residue (a lane is an edge iff it is
// zero).
template <int K, class V, bool RC = false>
IGC_HD inline V lvt_residue(const V (&m)[LvShape<K>::WD])
{
    constexpr int C = LvShape<K>::C;
    constexpr int WD = LvShape<K>::WD;
    if constexpr (K == 0) {
        return m[0];
    } else {
        V T[WD], S[WD]; // masks of level e-1, and the same plus one
IGC_UNROLL
        for (int t = 0; t < WD; ++t) {
            T[t] = V{}; // placeholders; only in-range entries are read
            S[t] = V{};
        }
        {
            const V x = m[C];
            T[C] = ~(x - 1) | x;
            S[C] = (V)(T[C] << 1);
        }
IGC_UNROLL
        for (int e = 1; e < K; ++e) {
            const int r = (e < K - e) ? e : (K - e);                    // range at level e
            const int rp = (e - 1 < K - e + 1) ? (e - 1) : (K - e + 1); // range at level e-1
            V NT[WD];
IGC_UNROLL
            for (int d = -C; d <= C; ++d) {
                NT[d + C] = T[d + C];
                if (d < -r || d > r) continue;
                V x = m[d + C];
                if (d >= -rp && d <= rp) x &= S[d + C];
                if (d - 1 >= -rp && d - 1 <= rp) x &= RC ? S[d - 1 + C] : T[d - 1 + C];
                if (d + 1 >= -rp && d + 1 <= rp) x &= RC ? T[d + 1 + C] : S[d + 1 + C];
                NT[d + C] = ~(x - 1) | x;
            }
IGC_UNROLL
            for (int t = 0; t < WD; ++t) {
                T[t] = NT[t];
                S[t] = (V)(NT[t] << 1);
            }
        }
        V L = S[C];
        if constexpr (K >= 2) {
            if constexpr (RC) L = L & S[C - 1] & T[C + 1];
            else L = L & T[C - 1] & S[C + 1];
        }
        return m[C] & L;
    }
}

// Row side: the row word's planes shifted for each diagonal, uniform across a
// wave (scalar registers in the kernels).
template <int K, class V>
struct LvtRow {
    V lo[LvShape<K>::WD];
    V hi[LvShape<K>::WD];
};

template <int K, class V, class W>
IGC_HD inline void lvt_row(W a, LvtRow<K, V> &r)
{
    constexpr int C = LvShape<K>::C;
IGC_UNROLL
    for (int t = 0; t < LvShape<K>::WD; ++t) row_planes(a, t - C, r.lo[t], r.hi[t]);
}

// Row coordinates on 16-bit lanes: the column side, each lane's two columns
// shifted for every diagonal (B << d, B >> -d within each 16-bit half; bits
// shifted past bit 15 or in at either end lie outside the diagonal's range).
template <int K, class V>
struct LvtCol {
    V lo[LvShape<K>::WD];
    V hi[LvShape<K>::WD];
};

template <int K, class V>
IGC_HD inline void lvt_col(V blo, V bhi, LvtCol<K, V> &c)
{
    constexpr int C = LvShape<K>::C;
IGC_UNROLL
    for (int t = 0; t < LvShape<K>::WD; ++t) {
        const int d = t - C;
        c.lo[t] = d >= 0 ? (V)(blo << d) : (V)(blo >> (-d));
        c.hi[t] = d >= 0 ? (V)(bhi << d) : (V)(bhi >> (-d));
    }
}

// The row's planes, unshifted, each copied into both halves of a register:
// lo = the word's low half twice, hi = its high half twice.
IGC_HD inline void lvt_row_rc(uint32_t a, u16x2 &lo, u16x2 &hi)
{
    const uint32_t l2 = (a << 16) | (a & 0xFFFFu);
    const uint32_t h2 = (a & 0xFFFF0000u) | (a >> 16);
    lo[0] = (uint16_t)l2;
    lo[1] = (uint16_t)(l2 >> 16);
    hi[0] = (uint16_t)h2;
    hi[1] = (uint16_t)(h2 >> 16);
}

// Edge lanes of the row (planes alo, ahi) against the columns c, in row
// coordinates.
template <int K, class V>
IGC_HD inline uint32_t lvt_test_rc(V alo, V ahi, const LvtCol<K, V> &c)
{
    V m[LvShape<K>::WD];
IGC_UNROLL
    for (int t = 0; t < LvShape<K>::WD; ++t) {
        m[t] = (alo ^ c.lo[t]) | (ahi ^ c.hi[t]);
#if defined(__HIP_DEVICE_COMPILE__)
        asm volatile("" : "+v"(m[t]));
#endif
    }
    return zero_lanes(lvt_residue<K, V, true>(m));
}

// Edge lanes (bit h: the pair in lane h has Levenshtein <= K).
template <int K, class V>
IGC_HD inline uint32_t lvt_test(const LvtRow<K, V> &r, V blo, V bhi)
{
    V m[LvShape<K>::WD];
IGC_UNROLL
    for (int t = 0; t < LvShape<K>::WD; ++t) {
        m[t] = (r.lo[t] ^ blo) | (r.hi[t] ^ bhi);
#if defined(__HIP_DEVICE_COMPILE__)
        // keep each mask in one register: without this the compiler folds the
        // OR into later ANDs and recomputes one plane's XOR per use
        asm volatile("" : "+v"(m[t]));
#endif
    }
    return zero_lanes(lvt_residue<K, V>(m));
}

// ---------------------------------------------------------------------------
// Bit-parallel Myers (Hyyro's formulation of Myers 1999): Levenshtein(T, P)
// for a text T and patterns P of one length n.  One step per text symbol; the
// match mask of text symbol T[j] is the set of pattern positions p with
// P_lo[p] == T_lo[j] and P_hi[p] == T_hi[j].  The vertical deltas vp/vn of the
// last column give D[n][n] = n + |vp| - |vn| over bits 0..n-1.  Bits at and
// above n hold padding; carries and shifts only move upward within a lane, so
// they never reach the counted bits.  T is the row (uniform across a wave), so
// its symbol bits are scalar work.  Returns the edge lanes (distance <= K).
template <class V>
IGC_HD inline uint32_t myers_lanes(uint32_t tlo, uint32_t thi, V plo, V phi, uint32_t n, int K)
{
    V vp = splat_lane<V>(0xFFFFFFFFu), vn = splat_lane<V>(0u);
    const V one = splat_lane<V>(1u);
    for (uint32_t j = 0; j < n; ++j) {
        const V s0 = splat_lane<V>(0u - ((tlo >> j) & 1u));
        const V s1 = splat_lane<V>(0u - ((thi >> j) & 1u));
        const V eq = ~((plo ^ s0) | (phi ^ s1));
        const V d0 = (((eq & vp) + vp) ^ vp) | eq | vn;
        const V hp = vn | ~(d0 | vp);
        const V hn = vp & d0;
        const V x = (V)(hp << 1) | one;
        vn = x & d0;
        vp = (V)(hn << 1) | ~(x | d0);
    }
    const uint32_t mask = n >= 32 ? ~0u : ((1u << n) - 1u);
    uint32_t out = 0;
IGC_UNROLL
    for (int h = 0; h < LaneCount<V>::value; ++h) {
        const int dist = (int)n + __builtin_popcount(lane_bits(vp, h) & mask) -
                         __builtin_popcount(lane_bits(vn, h) & mask);
        out |= (uint32_t)(dist <= K) << h;
    }
    return out;
}

// Row planes of the text for myers_lanes.
IGC_HD inline void text_planes(uint32_t a, uint32_t &lo, uint32_t &hi)
{
    lo = a & 0xFFFFu;
    hi = a >> 16;
}
IGC_HD inline void text_planes(uint64_t a, uint32_t &lo, uint32_t &hi)
{
    lo = (uint32_t)a;
    hi = (uint32_t)(a >> 32);
}

// ---------------------------------------------------------------------------
// Binary words (one plane, n <= 64): the same two tests on uint64_t lanes.
// Diagonal d's mismatch mask is (A >> d) ^ B (A << -d for d < 0); the
// thermometer argument above holds unchanged on a 64-bit lane, and at n = 64
// the lane's end is the strings' end.

template <int K>
struct LvtRowBin {
    uint64_t w[LvShape<K>::WD];
};

template <int K>
IGC_HD inline void lvt_row_bin(uint64_t a, LvtRowBin<K> &r)
{
    constexpr int C = LvShape<K>::C;
IGC_UNROLL
    for (int t = 0; t < LvShape<K>::WD; ++t) {
        const int s = t - C;
        r.w[t] = s >= 0 ? (a >> s) : (a << (-s));
    }
}

// 1 when the pair (row, b) has Levenshtein <= K
template <int K>
IGC_HD inline uint32_t lvt_test_bin(const LvtRowBin<K> &r, uint64_t b)
{
    uint64_t m[LvShape<K>::WD];
IGC_UNROLL
    for (int t = 0; t < LvShape<K>::WD; ++t) {
        m[t] = r.w[t] ^ b;
#if defined(__HIP_DEVICE_COMPILE__)
        asm volatile("" : "+v"(m[t]));
#endif
    }
    return zero_lanes(lvt_residue<K, uint64_t>(m));
}

// Myers with text t (the row) and pattern p (the column): the match mask of
// text symbol t[j] is p where t[j] = 1 and ~p where t[j] = 0.  Returns 1 when
// the distance is <= K.
IGC_HD inline uint32_t myers_bin(uint64_t t, uint64_t p, uint32_t n, int K)
{
    uint64_t vp = ~0ull, vn = 0ull;
    for (uint32_t j = 0; j < n; ++j) {
        const uint64_t s = 0ull - ((t >> j) & 1ull);
        const uint64_t eq = ~(p ^ s);
        const uint64_t d0 = (((eq & vp) + vp) ^ vp) | eq | vn;
        const uint64_t hp = vn | ~(d0 | vp);
        const uint64_t hn = vp & d0;
        const uint64_t x = (hp << 1) | 1ull;
        vn = x & d0;
        vp = (hn << 1) | ~(x | d0);
    }
    const uint64_t mask = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
    const int dist = (int)n + __builtin_popcountll(vp & mask) - __builtin_popcountll(vn & mask);
    return dist <= K ? 1u : 0u;
}

// ---------------------------------------------------------------------------
// Which predicate a (length, threshold) pair uses: the one with the lower
// measured time per pair, and no test at all when K >= n.  Word classes:
// 0 = two pairs per register (u16x2, n <= 16), 1 = one pair per 32-bit lane
// (17 <= n <= 32), 2 = binary words on 64-bit lanes (33 <= n <= 64).
enum class PredKind : int { LV = 0, MYERS = 1, ALL = 2 };

IGC_HD inline int word_class_of(uint32_t n) { return n <= 16 ? 0 : (n <= 32 ? 1 : 2); }

// Femtoseconds per pair of the count-only pair kernel at its fastest tile
// shape on MI355X, N = 65536 random strings, each predicate forced in turn:
// Landau-Vishkin per K for each word class (its cost does not depend on n) in
// the form the kernels run (16-bit lanes: row coordinates up to K = 3, column
// coordinates above), Myers a least-squares line in n per word class (one
// step per symbol).
IGC_HD inline uint32_t lv_fs(int wc, int K)
{
    const uint32_t t[3][kLvLaunchMaxK + 1] = {
        {92, 120, 276, 422, 722, 997, 1491, 2010, 2572, 3121, 3769, 4429, 5257, 6005},
        {150, 192, 493, 774, 1243, 1844, 2625, 3592, 4502, 5708, 7062, 8361, 9990, 11455},
        {152, 218, 672, 1150, 1986, 2997, 4239, 5559, 7189, 8921, 11123, 13123, 15748, 18382}};
    return t[wc][K];
}

IGC_HD inline uint32_t myers_fs(int wc, uint32_t n)
{
    const uint32_t a[3] = {386, 250, 259}, b[3] = {145, 241, 515};
    return a[wc] + b[wc] * n;
}

IGC_HD inline PredKind pred_kind(uint32_t n, int K)
{
    if (K >= kAllPairs || K >= (int)n) return PredKind::ALL;
    const int wc = word_class_of(n);
    if (K <= kLvLaunchMaxK && lv_fs(wc, K) <= myers_fs(wc, n)) return PredKind::LV;
    return PredKind::MYERS;
}

} // namespace igc
