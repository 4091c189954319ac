// igc_lv.hpp -- the pair predicate of the conflict-graph kernels, written so
// the same code compiles for gfx950 (inside the HIP kernels) and for the host
// (where the tests check it exhaustively against the oracle).
//
// Packed word of a candidate (n <= 16): bit p = sym[p] & 1, bit 16+p = sym[p] >> 1,
// all other bits 0.  lv_edge<K> decides Levenshtein(A, B) <= K by
// Landau-Vishkin furthest-reaching diagonals over mismatch bitmasks; see the
// header of igc_kernels.hpp for the derivation.
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

constexpr int kAllPairs = 16; // K value meaning "every pair is an edge" (K >= n)

IGC_HD inline uint32_t pack_word(const uint8_t *s, uint32_t n)
{
    uint32_t w = 0;
    for (uint32_t p = 0; p < n; ++p) {
        w |= (uint32_t)(s[p] & 1u) << p;
        w |= (uint32_t)((s[p] >> 1) & 1u) << (16 + p);
    }
    return w;
}

// Mismatch mask of diagonal d for row word `sa` (= A shifted by d) and column
// word b: bit j (j < 16) = symbol mismatch, bits 16..31 forced to 1.
IGC_HD inline uint32_t mism_mask(uint32_t sa, uint32_t b)
{
    const uint32_t x = sa ^ b;
    return x | (x >> 16) | 0xFFFF0000u;
}

template <int K>
struct LvShape {
    static constexpr int C = (K >= kAllPairs) ? 0 : K / 2; // band half-width
    static constexpr int WD = 2 * C + 1;                   // diagonals used
};

// Row word shifted for each diagonal d = t - C (uniform across the wave).
template <int K>
IGC_HD inline void row_shifts(uint32_t a, uint32_t (&sa)[LvShape<K>::WD])
{
    constexpr int C = LvShape<K>::C;
IGC_UNROLL
    for (int t = 0; t < LvShape<K>::WD; ++t) {
        const int d = t - C;
        sa[t] = d >= 0 ? (a >> d) : (a << (-d));
    }
}

// Levenshtein(A,B) <= K for the packed words (see the file header).
template <int K>
IGC_HD inline bool lv_edge(const uint32_t (&sa)[LvShape<K>::WD], uint32_t b)
{
    if constexpr (K >= kAllPairs) {
        return true;
    } else if constexpr (K == 0) {
        const uint32_t x = sa[0] ^ b;
        return ((x | (x >> 16)) & 0xFFFFu) == 0u;
    } else {
        constexpr int C = LvShape<K>::C;
        constexpr int WD = LvShape<K>::WD;
        uint32_t m[WD];
        uint32_t q0 = 0;
IGC_UNROLL
        for (int t = 0; t < WD; ++t) {
            const uint32_t x = sa[t] ^ b;
            const uint32_t h = x >> 16;
            m[t] = x | h | 0xFFFF0000u;
            if (t == C) q0 = (x & 0xFFFFu) | h;
        }
        uint32_t fr[WD], fv[WD]; // FR and FR+1 of the previous level
IGC_UNROLL
        for (int t = 0; t < WD; ++t) {
            fr[t] = 0u;
            fv[t] = 0u;
        }
        fr[C] = (uint32_t)__builtin_ctz(m[C]);
        fv[C] = fr[C] + 1u;
IGC_UNROLL
        for (int e = 1; e < K; ++e) {
            const int r = (e < K - e) ? e : (K - e);              // range at level e
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
                const uint32_t s = L + (uint32_t)__builtin_ctz(m[d + C] >> L);
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
        return (q0 >> L) == 0u;
    }
}

} // namespace igc
