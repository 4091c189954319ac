// Host build of the GPU pair predicates (indexgen_amd/csrc/igc_lv.hpp) for the
// CPU tests: the same template code the kernels run, compiled with g++.
#include "../indexgen_amd/csrc/igc_lv.hpp"

#include <stddef.h>
#include <stdint.h>

using namespace igc;

template <int K>
static void run32(const uint8_t *A, const uint8_t *B, int M, int n, uint8_t *out)
{
    for (int r = 0; r < M; ++r) {
        uint32_t sa[LvShape<K>::WD];
        row_shifts<K>(pack_word(A + (size_t)r * n, n), sa);
        out[r] = lv_edge<K>(sa, pack_word(B + (size_t)r * n, n)) ? 1 : 0;
    }
}

template <int K>
static void run64(const uint8_t *A, const uint8_t *B, int M, int n, uint8_t *out)
{
    for (int r = 0; r < M; ++r) {
        uint64_t sa[LvShape<K>::WD];
        row_shifts<K>(pack_word64(A + (size_t)r * n, n), sa);
        out[r] = lv_edge<K>(sa, pack_word64(B + (size_t)r * n, n)) ? 1 : 0;
    }
}

// out[r] = (Levenshtein(A[r], B[r]) <= K) by Landau-Vishkin, 0 <= K <= 15, on
// 32-bit words (n <= 16) or 64-bit words (wide, n <= 32)
extern "C" int lv_pairs(const uint8_t *A, const uint8_t *B, int M, int n, int K, int wide, uint8_t *out)
{
    if (wide ? n > 32 : n > 16) return -1;
    switch (K) {
#define C_(k) case k: if (wide) run64<k>(A, B, M, n, out); else run32<k>(A, B, M, n, out); return 0;
        C_(0) C_(1) C_(2) C_(3) C_(4) C_(5) C_(6) C_(7) C_(8) C_(9) C_(10) C_(11) C_(12) C_(13) C_(14) C_(15)
#undef C_
    }
    return -1;
}

// the same by the Myers predicate (A is the text, B the pattern)
extern "C" int myers_pairs(const uint8_t *A, const uint8_t *B, int M, int n, int K, int wide, uint8_t *out)
{
    if (wide ? n > 32 : n > 16) return -1;
    for (int r = 0; r < M; ++r) {
        const uint8_t *a = A + (size_t)r * n, *b = B + (size_t)r * n;
        const bool e = wide ? myers_edge<uint64_t>(pack_word64(a, n), pack_word64(b, n), (uint32_t)n, K)
                            : myers_edge<uint32_t>(pack_word(a, n), pack_word(b, n), (uint32_t)n, K);
        out[r] = e ? 1 : 0;
    }
    return 0;
}

// the predicate the kernels pick for (n, K): 0 LV32, 1 LV64, 2 Myers, 3 every pair
extern "C" int pred_kind_of(int n, int K) { return (int)pred_kind((uint32_t)n, K); }
extern "C" int lv_launch_max_k(void) { return kLvLaunchMaxK; }
