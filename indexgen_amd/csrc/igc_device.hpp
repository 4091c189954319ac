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
#include <vector>

#define HIP_TRY(expr)                                                                                         \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return ::igc::set_error(e_ == hipErrorOutOfMemory ? IGC_ERR_OOM : IGC_ERR_DEVICE,                 \
                                    "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

namespace igc {

// Device-memory budget shared by several contexts (the server's pipelines).
// Each context draws from it through its own MemSlot: `used` is what the
// slot's buffers hold, `charged` what the slot has reserved (>= used).  The
// sum of all charges never exceeds the limit, so neither do the buffers.
struct MemPool {
    std::mutex mu;
    uint64_t limit = ~0ull;
    uint64_t charged = 0; // sum over slots
    uint64_t peak = 0;    // highest sum of charges
};

struct MemSlot {
    MemPool *pool = nullptr;
    uint64_t used = 0, charged = 0;

    // Reserves at least `bytes` (and at least what the slot holds); false when
    // the pool's limit would be exceeded.
    bool reserve(uint64_t bytes)
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        const uint64_t want = std::max(bytes, used);
        if (want <= charged) return true;
        if (pool->charged + (want - charged) > pool->limit) return false;
        pool->charged += want - charged;
        charged = want;
        pool->peak = std::max(pool->peak, pool->charged);
        return true;
    }
    // A buffer of the slot grows by delta bytes: within the reservation, or
    // by extending it when the pool allows.
    bool grow(uint64_t delta)
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        if (used + delta > charged) {
            const uint64_t need = used + delta - charged;
            if (pool->charged + need > pool->limit) return false;
            pool->charged += need;
            charged += need;
            pool->peak = std::max(pool->peak, pool->charged);
        }
        used += delta;
        return true;
    }
    void shrink(uint64_t delta)
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        used -= std::min(used, delta);
    }
    // Gives back the reservation above what the buffers hold.
    void settle()
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        pool->charged -= charged - used;
        charged = used;
    }
};

// grow-only device buffer; with a slot, every allocation is charged to it
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    MemSlot *slot = nullptr;
    int ensure(size_t want)
    {
        if (want <= bytes && p) return IGC_OK;
        release();
        const size_t sz = std::max<size_t>(want, 256);
        if (slot && !slot->grow(sz))
            return set_error(IGC_ERR_OOM, "device buffer of %zu bytes exceeds the HBM budget (%llu of %llu bytes "
                                          "reserved)", sz, (unsigned long long)slot->pool->charged,
                             (unsigned long long)slot->pool->limit);
        const hipError_t e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            p = nullptr;
            if (slot) slot->shrink(sz);
            return set_error(IGC_ERR_OOM, "hipMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
        }
        bytes = sz;
        return IGC_OK;
    }
    void release()
    {
        if (p) {
            (void)hipFree(p);
            if (slot) slot->shrink(bytes);
        }
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
// First edge buffer (entries) for N strings of length n at threshold K: twice
// the density estimate, or kUnknownDensityCap entries where there is none (an
// overflow is regrown to the exact count and the pairs replayed).
constexpr uint64_t kUnknownDensityCap = 1ull << 26;
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
    // first edge buffer of a build, in entries (0: sized by the density
    // estimate); a smaller one is regrown to the exact count and the pairs replayed
    uint64_t first_edge_cap = 0;
    // the refinement loop replays its captured round (igc_ctx_set_graph_replay)
    bool graph_replay = true;
    // charge account of the context's device buffers when it draws from a
    // shared budget (the server's pipelines); no pool = not charged
    igc::MemSlot mem;

    // every device buffer the context owns directly (not the K-means buffers)
    template <class F> void for_each_devbuf(F f)
    {
        for (igc::DevBuf *b : {&syms, &packed, &edges, &misc, &row_ptr, &col, &bsum, &sets, &bounds, &srcs, &keys[0],
                               &keys[1], &rcount, &roffset, &segmeta, &nbase, &feat, &feat_s, &nkey, &perm, &packed_s,
                               &fitems, &fbias, &fcount, &fsets, &gen_out, &gen_cnt, &gen_off, &gen_rows})
            f(*b);
        for (igc::DevBuf &b : corder) f(b);
    }
};

namespace igc {
// Sizes of the segmented CSR form (a split pass on the row's high hb bits,
// then 32-bit passes over rowbits + dbits low bits) for an edge capacity cap.
struct SegPlan {
    uint32_t N = 0, dbits = 0, rowbits = 0, hb = 0, nseg = 0, lb = 0, passes = 0, bits = 0, Lmax = 0;
    bool upper = false;
    uint64_t cap = 0, items_cap = 0, tiles64 = 0, ub = 0;
    size_t slots = 0;
    size_t bytes_keys0 = 0, bytes_keys1 = 0, bytes_rcount = 0, bytes_roffset = 0, bytes_bsum = 0, bytes_segmeta = 0;
};
// Device buffers of the segmented form: the CSR (row_ptr N+1, col items_cap)
// and its temporaries, each at least the plan's size.
struct SegBufs {
    uint64_t *row_ptr;
    uint32_t *col;
    uint32_t *keys0, *keys1, *rcount;
    uint64_t *roffset, *bsum;
    SegMeta *segmeta;
};
int seg_plan(uint32_t N, uint64_t cap, bool upper, uint32_t dbits, uint32_t rowbits, uint32_t hb, SegPlan &P);
// Asynchronous on s; reads the edge count from counter (misc words 0..1) and
// uses misc word 4.  Every buffer is written before it is read (no clearing).
void launch_csr_seg_into(hipStream_t s, const SegPlan &P, const SegBufs &B, const uint64_t *edges,
                         unsigned long long *counter, uint32_t *misc, uint32_t *launches, const uint32_t *base);
// Builds the CSR of the edge list in ctx->edges (count in counter, capacity cap)
// over N nodes into ctx->row_ptr / ctx->col; asynchronous on ctx->stream.
int launch_csr(igc_ctx *ctx, uint32_t N, uint64_t cap, unsigned long long *counter, uint32_t *misc,
               uint32_t *launches, bool upper = false, uint32_t maxset = 0, const uint32_t *base = nullptr);
Shape pick_shape(igc_ctx *ctx, uint32_t n, int K, uint32_t N);

int filter_sets(igc_ctx *ctx, const uint8_t *d_syms, const void *packed, void *packed_s, uint2 *feat_s,
                uint32_t *perm, uint4 *items, uint32_t *bias, uint32_t *count, std::vector<igc::FilterSet> &fs);

// initial_edge_cap, or the context's forced first capacity (igc_ctx_set_edge_cap)
uint64_t first_edge_cap(const igc_ctx *ctx, uint32_t N, uint32_t n, int K);

int kmeans_attach(igc_ctx *ctx);
void kmeans_release(igc_ctx *ctx);
int kmeans_device_locked(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n, uint32_t k, uint32_t iters,
                         uint64_t seed, bool use_graph, float *ms_out);
int32_t *kmeans_assign_ptr(igc_ctx *ctx);
uint32_t *kmeans_flags_ptr(igc_ctx *ctx); // kMiscWords words; [2] bad symbol flag
// The refinement loop's K-means: buffers for up to Nmax points (kept in
// place, so captured launches stay valid), the seeds of a round (host_seeds
// must live until the stream has copied them), and the whole K-means --
// pack, seeds, `iters` Lloyd iterations, final assignment -- as kernels only,
// with grids for Nmax and the round's count read from dN (capturable).
int kmeans_prepare_locked(igc_ctx *ctx, uint32_t Nmax, uint32_t n, uint32_t k);
int kmeans_seed_locked(igc_ctx *ctx, uint32_t N, uint32_t k, uint64_t seed, std::vector<uint32_t> &host_seeds);
void kmeans_launch_locked(igc_ctx *ctx, const uint8_t *d_syms, uint32_t Nmax, const uint32_t *dN, uint32_t n,
                          uint32_t k, uint32_t iters);

// Looks up the device address of the no-allocation build's scratch array on
// `device` (the current device) once; igc_ctx_create calls it (igc_noalloc.hip).
int scratch_attach(int device);

// Charges every device buffer of ctx (present and future) to a slot of a
// shared pool; buffers already held are charged at once.
void ctx_attach_pool(igc_ctx *ctx, MemPool *pool);
// Frees every device buffer of ctx (they are regrown on the next build) and
// gives back the slot's reservation; the caller holds ctx->mu.
void ctx_trim_locked(igc_ctx *ctx);

// Device-to-host streaming through a ring of pinned staging chunks consumed on
// worker threads (igc_runtime.hip); the caller holds the context lock and the
// source is complete (the context's stream has reached it).
constexpr size_t kD2HChunk = 4u << 20;
constexpr uint32_t kD2HThreads = 8;
using D2HConsumer = std::function<void(const uint8_t *chunk, size_t offset, size_t bytes)>;
int d2h_stream(igc_ctx *ctx, const void *src, size_t bytes, const D2HConsumer &consume);

// igc_graph_build_batch without taking the context lock (caller holds it)
int batch_build_host_locked(igc_ctx *ctx, const igc_set *sets, uint32_t count, igc_csr *outs);
// The same with the sets' symbols already concatenated in device memory at
// d_syms, in set order (sets[s].syms is not read; it must be non-null when N > 0)
// The refinement loop (igc_refine.hip) on a locked context.
int refine_locked(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist, uint32_t k,
                  uint32_t kmeans_iters, uint32_t max_rounds, uint32_t converge_n, uint64_t seed,
                  uint32_t host_threads, bool use_graph, uint32_t *out_idx, uint32_t *out_count,
                  igc_refine_stats *stats);
int batch_build_dev_syms_locked(igc_ctx *ctx, const uint8_t *d_syms, const igc_set *sets, uint32_t count,
                                igc_csr *outs);
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
