// igc_device.hpp -- device-side plumbing shared by the HIP translation units of
// libigc.so: the context, grow-only device buffers and the launch helpers.
#pragma once
#include "igc_internal.hpp"
#include "igc_csr.hpp"
#include "igc_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <mutex>

#define HIP_TRY(expr)                                                                                         \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return ::igc::set_error(e_ == hipErrorOutOfMemory ? IGC_ERR_OOM : IGC_ERR_DEVICE,                 \
                                    "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

namespace igc {

// grow-only device buffer
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want)
    {
        if (want <= bytes && p) return IGC_OK;
        release();
        const size_t sz = std::max<size_t>(want, 256);
        const hipError_t e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            p = nullptr;
            return set_error(IGC_ERR_OOM, "hipMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
        }
        bytes = sz;
        return IGC_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// grow-only pinned host buffer (fast device-to-host copies of the CSR)
struct HostBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want)
    {
        if (want <= bytes && p) return IGC_OK;
        release();
        const size_t sz = std::max<size_t>(want, 256);
        const hipError_t e = hipHostMalloc(&p, sz, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return set_error(IGC_ERR_OOM, "hipHostMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
        }
        bytes = sz;
        return IGC_OK;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// misc buffer layout (uint32 words): [0..1] edge counter (u64), [2] bad symbol
// flag, [4] tiles of the segmented CSR form; the rest is spare
constexpr size_t kMiscWords = 8;

// Tile shape of the single-set kernel: column slots per lane (each slot one
// register of lane_cols(n) columns) and rows per tile W/q.
struct Shape {
    int cpl;
    uint32_t q;
};

// Tile geometry for a workgroup of 4 waves with cols_per_lane columns per lane
// (slots per lane x the policy's lanes per register) and W/q rows per tile.
TileGeom choose_geometry(uint32_t N, int cols_per_lane, uint32_t q);
uint64_t tile_count(const TileGeom &G);
Shape default_shape(uint32_t N);
uint64_t initial_edge_cap(uint32_t N, uint32_t n, int K);
int effective_k(uint32_t n, uint32_t edit_dist);
uint32_t grid_for(uint64_t work, uint32_t per_block, uint32_t cap_blocks);

// Packed words are 32-bit for n <= 16, 64-bit (two planes) for 17 <= n <= 32,
// and 64-bit binary words (one plane, symbols 0/1) for 33 <= n <= 64.
inline bool wide_words(uint32_t n) { return n > 16; }
inline bool binary_words(uint32_t n) { return n > IGC_MAX_QUATERNARY_LEN; }
// 0: 32-bit words, 1: 64-bit two-plane words, 2: 64-bit binary words
inline int word_class(uint32_t n) { return word_class_of(n); }
static_assert(IGC_MAX_QUATERNARY_LEN == 32, "word_class_of splits at n = 32");
inline size_t word_bytes(uint32_t n) { return wide_words(n) ? 8 : 4; }
// Columns a pair test covers per register: two 16-bit lanes for n <= 16.
inline int lane_cols(uint32_t n) { return wide_words(n) ? 1 : 2; }

// Launches the pair kernel of the predicate pred_kind(n, K) picks, with cpl
// columns per lane, over `tiles` tiles of geometry G; `packed` holds words of
// the width of n.
int launch_pairs(PredKind pk, uint32_t n, int K, int cpl, bool emit, hipStream_t s, uint64_t tiles, const void *packed,
                 const TileGeom &G, uint32_t node_off, uint64_t *edges, uint64_t cap, unsigned long long *counter,
                 const RowFilter &f = RowFilter{});
int launch_pack(hipStream_t s, const uint8_t *d_syms, uint32_t N, uint32_t n, void *packed, uint32_t *bad);
// offsets[0..L] = exclusive scan of counts[0..L-1] (offsets[L] = total); bsum
// holds ceil(L / kScanTile) partial sums
void scan_counts(hipStream_t s, const uint32_t *counts, uint32_t L, uint64_t *bsum, uint64_t *offsets);
int pred_variant(PredKind pk, uint32_t n, int K);
constexpr int kPredVariants = 3 * (kLvLaunchMaxK + 1) + 6;

// Whether a set of N strings gets the row filter under prefilter `mode`: it
// needs a nontrivial threshold, and automatically only sets of at least
// IGC_PREFILTER_MIN_N strings (below it the sort's launches cost more than the
// rows it saves).  Binary lengths (n > 32) run unfiltered.
inline bool filter_pays(int mode, uint32_t N, uint32_t n, int K)
{
    if (mode == IGC_PREFILTER_OFF || N < 2 || K < 1 || binary_words(n) || pred_kind(n, K) == PredKind::ALL)
        return false;
    return mode == IGC_PREFILTER_ALWAYS || N >= IGC_PREFILTER_MIN_N;
}

// Error for a packing pass that flagged symbols: bit 0 a code above 3, bit 1 a
// code above 1 in a binary string (n > 32).
inline int bad_symbols(uint32_t flags)
{
    if (flags & 2u)
        return set_error(IGC_ERR_INVALID, "symbol code above 1 in a binary string (strings longer than %d take "
                                          "symbols 0/1 only)", IGC_MAX_QUATERNARY_LEN);
    return set_error(IGC_ERR_INVALID, "symbol code above 3 in input");
}

} // namespace igc

struct igc_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::mutex mu;
    igc::DevBuf syms, packed, edges, misc, row_ptr, col, bsum, sets, bounds, srcs;
    igc::DevBuf keys[2];         // radix-sort ping-pong buffers (2 entries per edge)
    igc::DevBuf rcount, roffset; // per-tile digit counts and their scan
    igc::DevBuf segmeta, nbase;  // segment table of the 32-bit CSR form; batch node -> first node of its set
    int csr_form = IGC_CSR_AUTO; // IGC_CSR_*
    // row filter: half counts (unsorted, sorted), bucket keys, permutation,
    // sorted packed words, window boxes and thresholds
    igc::DevBuf feat, feat_s, nkey, perm, packed_s, fitems, fbias, fcount, fsets;
    std::vector<igc::FilterSet> fs_host; // the last filter set-up's table (read by its copy)
    igc::DevBuf corder[IGC_MAX_LEN + 1]; // class order along the curve, per string length
    int prefilter = IGC_PREFILTER_AUTO; // exact row skipping (IGC_PREFILTER_*)
    int predicate = IGC_PRED_AUTO;      // forced pair predicate (IGC_PRED_*)
    igc::HostBuf host_rp, stage; // pinned: batch row pointers, staged copies
    igc::DevBuf gen_out, gen_cnt, gen_off, gen_rows; // candidate generators: output, per-block counts/offsets, rows
    uint64_t gen_serial = 0;                         // generations so far
    hipEvent_t ev[5] = {};
    std::vector<hipEvent_t> cev; // staging-ring events of d2h_stream
    // auxiliary streams (default priority) and their fork/join events: a batch's
    // predicate groups run concurrently (igc_batch.hip)
    static constexpr uint32_t kAuxStreams = 3;
    hipStream_t aux[kAuxStreams] = {};
    hipEvent_t aux_ev[kAuxStreams] = {};
    igc_timing last{};
    int cu_count = 256;
    // tile shapes chosen by the autotuner, keyed by (kernel variant, N bucket);
    // cpl 0 = not tuned
    igc::Shape tuned[igc::kPredVariants][8] = {};
    void *km = nullptr; // K-means buffers (igc_kmeans.hip)
};

namespace igc {
// Builds the CSR of the edge list in ctx->edges (count in counter, capacity cap)
// over N nodes into ctx->row_ptr / ctx->col; asynchronous on ctx->stream.
int launch_csr(igc_ctx *ctx, uint32_t N, uint64_t cap, unsigned long long *counter, uint32_t *misc,
               uint32_t *launches, bool upper = false, uint32_t maxset = 0, const uint32_t *base = nullptr);
Shape pick_shape(igc_ctx *ctx, uint32_t n, int K, uint32_t N);

int filter_sets(igc_ctx *ctx, const uint8_t *d_syms, const void *packed, void *packed_s, uint2 *feat_s,
                uint32_t *perm, uint4 *items, uint32_t *bias, uint32_t *count, std::vector<igc::FilterSet> &fs);

int kmeans_attach(igc_ctx *ctx);
void kmeans_release(igc_ctx *ctx);
int kmeans_device_locked(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n, uint32_t k, uint32_t iters,
                         uint64_t seed, bool use_graph, float *ms_out);
int32_t *kmeans_assign_ptr(igc_ctx *ctx);

// Device-to-host streaming through a ring of pinned staging chunks consumed on
// worker threads (igc_runtime.hip); the caller holds the context lock and the
// source is complete (the context's stream has reached it).
constexpr size_t kD2HChunk = 4u << 20;
constexpr uint32_t kD2HThreads = 8;
using D2HConsumer = std::function<void(const uint8_t *chunk, size_t offset, size_t bytes)>;
int d2h_stream(igc_ctx *ctx, const void *src, size_t bytes, const D2HConsumer &consume);

// igc_graph_build_batch without taking the context lock (caller holds it)
int batch_build_host_locked(igc_ctx *ctx, const igc_set *sets, uint32_t count, igc_csr *outs);
} // namespace igc

namespace igc {
// The predicate the context's kernels run for (n, K): pred_kind unless the
// context forces one (every pair stays every pair; Landau-Vishkin only where
// it is instantiated).
inline PredKind ctx_pred(const igc_ctx *ctx, uint32_t n, int K)
{
    const PredKind pk = pred_kind(n, K);
    if (pk == PredKind::ALL || ctx->predicate == IGC_PRED_AUTO) return pk;
    if (ctx->predicate == IGC_PRED_LANDAU_VISHKIN) return K <= kLvLaunchMaxK ? PredKind::LV : pk;
    return PredKind::MYERS;
}
} // namespace igc
