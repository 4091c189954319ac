// Host build of the GPU pair predicates (indexgen_amd/csrc/igc_lv.hpp) for the
// CPU tests: the same template code the kernels run, compiled with g++.
//
// Each predicate runs on the lane types and in the forms the kernels use: two
// pairs per call on 16-bit lanes (n <= 16; Landau-Vishkin in row coordinates
// and in column coordinates) or one pair on a 32-bit lane (wide, n <= 32;
// column coordinates).  For 16-bit lanes every pair (A[r], B[r]) is run in
// each lane, with an unrelated column string in the other lane, and by each
// form; out[r] is 2 when any two of these disagree.  Binary strings (n <= 64)
// run on 64-bit lanes.
#include "../indexgen_amd/csrc/igc_lv.hpp"

#include <stddef.h>
#include <stdint.h>

using namespace igc;

namespace {

struct LvPred {
    template <int K>
    static void run(const uint8_t *A, const uint8_t *B, int M, int n, int wide, uint8_t *out)
    {
        for (int r = 0; r < M; ++r) {
            const uint8_t *a = A + (size_t)r * n, *b = B + (size_t)r * n;
            const uint8_t *o = B + (size_t)((r + 1) % M) * n;
            if (wide) {
                LvtRow<K, uint32_t> row;
                lvt_row<K>(pack_word64(a, n), row);
                uint32_t lo, hi;
                col_planes(pack_word64(b, n), lo, hi);
                out[r] = (uint8_t)lvt_test<K, uint32_t>(row, lo, hi);
            } else {
                // both forms the kernels run on 16-bit lanes, each pair in both lanes
                u16x2 alo, ahi, lo, hi;
                lvt_row_rc(pack_word(a, n), alo, ahi);
                LvtRow<K, u16x2> row;
                lvt_row<K>(pack_word(a, n), row);
                LvtCol<K, u16x2> col;
                uint32_t e[4];
                col_planes(pack_word(b, n), pack_word(o, n), lo, hi);
                lvt_col<K>(lo, hi, col);
                e[0] = lvt_test_rc<K, u16x2>(alo, ahi, col) & 1u;
                e[1] = lvt_test<K, u16x2>(row, lo, hi) & 1u;
                col_planes(pack_word(o, n), pack_word(b, n), lo, hi);
                lvt_col<K>(lo, hi, col);
                e[2] = (lvt_test_rc<K, u16x2>(alo, ahi, col) >> 1) & 1u;
                e[3] = (lvt_test<K, u16x2>(row, lo, hi) >> 1) & 1u;
                out[r] = (e[0] == e[1] && e[0] == e[2] && e[0] == e[3]) ? (uint8_t)e[0] : 2;
            }
        }
    }
};

} // namespace

// out[r] = (Levenshtein(A[r], B[r]) <= K) by Landau-Vishkin, 0 <= K <= 15
extern "C" int lv_pairs(const uint8_t *A, const uint8_t *B, int M, int n, int K, int wide, uint8_t *out)
{
    if (wide ? n > 32 : n > 16) return -1;
    switch (K) {
#define C_(k) case k: LvPred::run<k>(A, B, M, n, wide, out); return 0;
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
        const uint8_t *o = B + (size_t)((r + 1) % M) * n;
        uint32_t tlo, thi;
        if (wide) {
            text_planes(pack_word64(a, n), tlo, thi);
            uint32_t lo, hi;
            col_planes(pack_word64(b, n), lo, hi);
            out[r] = (uint8_t)myers_lanes<uint32_t>(tlo, thi, lo, hi, (uint32_t)n, K);
        } else {
            text_planes(pack_word(a, n), tlo, thi);
            u16x2 lo, hi;
            col_planes(pack_word(b, n), pack_word(o, n), lo, hi);
            const uint32_t e0 = myers_lanes<u16x2>(tlo, thi, lo, hi, (uint32_t)n, K) & 1u;
            col_planes(pack_word(o, n), pack_word(b, n), lo, hi);
            const uint32_t e1 = (myers_lanes<u16x2>(tlo, thi, lo, hi, (uint32_t)n, K) >> 1) & 1u;
            out[r] = e0 == e1 ? (uint8_t)e0 : 2;
        }
    }
    return 0;
}

namespace {
struct LvBin {
    template <int K>
    static void run(const uint8_t *A, const uint8_t *B, int M, int n, uint8_t *out)
    {
        for (int r = 0; r < M; ++r) {
            LvtRowBin<K> row;
            lvt_row_bin<K>(pack_bin64(A + (size_t)r * n, n), row);
            out[r] = (uint8_t)lvt_test_bin<K>(row, pack_bin64(B + (size_t)r * n, n));
        }
    }
};
} // namespace

// binary strings (symbols 0/1, n <= 64) on 64-bit lanes: Landau-Vishkin
extern "C" int lv_pairs_bin(const uint8_t *A, const uint8_t *B, int M, int n, int K, uint8_t *out)
{
    if (n > 64) return -1;
    switch (K) {
#define C_(k) case k: LvBin::run<k>(A, B, M, n, out); return 0;
        C_(0) C_(1) C_(2) C_(3) C_(4) C_(5) C_(6) C_(7) C_(8) C_(9) C_(10) C_(11) C_(12) C_(13) C_(14) C_(15)
#undef C_
    }
    return -1;
}

// ... and Myers (A is the text, B the pattern)
extern "C" int myers_pairs_bin(const uint8_t *A, const uint8_t *B, int M, int n, int K, uint8_t *out)
{
    if (n > 64) return -1;
    for (int r = 0; r < M; ++r)
        out[r] = (uint8_t)myers_bin(pack_bin64(A + (size_t)r * n, n), pack_bin64(B + (size_t)r * n, n), (uint32_t)n, K);
    return 0;
}

// the predicate the kernels pick for (n, K): 0 Landau-Vishkin, 1 Myers, 2 every pair
extern "C" int pred_kind_of(int n, int K) { return (int)pred_kind((uint32_t)n, K); }
extern "C" int lv_launch_max_k(void) { return kLvLaunchMaxK; }
