// Host build of the GPU pair predicate (indexgen_amd/csrc/igc_lv.hpp) for the
// CPU tests: the same template code the kernels run, compiled with g++.
#include "../indexgen_amd/csrc/igc_lv.hpp"

#include <stddef.h>
#include <stdint.h>

using namespace igc;

template <int K>
static void run(const uint8_t *A, const uint8_t *B, int M, int n, uint8_t *out)
{
    for (int r = 0; r < M; ++r) {
        uint32_t sa[LvShape<K>::WD];
        row_shifts<K>(pack_word(A + (size_t)r * n, n), sa);
        out[r] = lv_edge<K>(sa, pack_word(B + (size_t)r * n, n)) ? 1 : 0;
    }
}

// out[r] = (Levenshtein(A[r], B[r]) <= K) as the kernel computes it, 0 <= K <= 16
extern "C" int lv_pairs(const uint8_t *A, const uint8_t *B, int M, int n, int K, uint8_t *out)
{
    switch (K) {
#define C_(k) case k: run<k>(A, B, M, n, out); return 0;
        C_(0) C_(1) C_(2) C_(3) C_(4) C_(5) C_(6) C_(7) C_(8) C_(9) C_(10) C_(11) C_(12) C_(13) C_(14) C_(15) C_(16)
#undef C_
    }
    return -1;
}
