// igc_noalloc.hip -- the single-set graph build for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// igc_graph_build_device_noalloc is igc_graph_build_device with the CSR
// written into caller buffers and every temporary -- the packed words, the
// edge list, the edge counter and the radix sort's keys, digit counts and
// segment table -- carved from one file-scope device array, `scratch`, of
// kScratchBytes.  Each GPU the library runs on holds its own copy of the array
// until the process ends.  The host never passes the array's name to a kernel:
// its device address is looked up once per device with hipGetSymbolAddress
// (at context creation, with the device current) and only that address is
// used.  The call makes no hipMalloc, hipFree or other allocation, host or
// device; it clears what must start at zero (the edge counter and flags) with
// a fill kernel.  The array keeps the last call's contents, so calls on one
// GPU must not overlap: the library serialises them with a per-device lock
// held until the call's work has finished.
//
// Limits (refused with IGC_ERR_INVALID and a message naming the size needed):
// the packed words (4 or 8 B per string), the edge list (8 B per edge) and the
// CSR sort's temporaries (about 16 B per edge) must fit in kScratchBytes, i.e.
// about 5 million edges; the row filter is not used (same graph either way).
#include "igc_device.hpp"

#include <mutex>

namespace igc {

constexpr size_t kScratchBytes = 128ull << 20;
constexpr int kScratchDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static uint8_t *g_scratch[kScratchDevices] = {};  // device address of `scratch`, per device
static std::mutex g_device_mu[kScratchDevices];   // one no-alloc call at a time per device

// Looks up (once) the device address of `scratch` on `device`, which must be
// the current device.
int scratch_attach(int device)
{
    if (device < 0 || device >= kScratchDevices)
        return set_error(IGC_ERR_INVALID, "device %d beyond the %d the no-allocation build supports", device,
                         kScratchDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return IGC_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(IGC_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                         hipGetErrorString(e));
    g_scratch[device] = static_cast<uint8_t *>(p);
    return IGC_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static int noalloc_locked(igc_ctx *ctx, uint8_t *scr, const uint8_t *d_syms, uint32_t N, uint32_t n,
                          uint32_t edit_dist, uint64_t *d_row_ptr, uint32_t *d_col, uint64_t col_cap,
                          uint64_t *n_edges)
{
    hipStream_t s = ctx->stream;
    igc_timing tm{};
    tm.pairs = (uint64_t)N * (N > 0 ? N - 1 : 0) / 2;
    *n_edges = 0;
    auto empty_csr = [&]() -> int {
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2 * ((uint64_t)N + 1), 256, 4096)), dim3(256), 0, s,
                           reinterpret_cast<uint32_t *>(d_row_ptr), 2 * ((uint64_t)N + 1), 0u);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
        return IGC_OK;
    };
    if (edit_dist == 0 || N < 2) {
        ctx->last = tm;
        return empty_csr();
    }
    const int K = effective_k(n, edit_dist);
    const Shape sh = pick_shape(ctx, n, K, N);
    const TileGeom G = choose_geometry(N, sh.cpl * lane_cols(n), sh.q);
    const uint64_t tiles = tile_count(G);
    // scratch: [flags and counter][packed words][edge list ... | CSR temporaries after the E edges]
    const size_t off_packed = 256;
    const size_t off_edges = align256(off_packed + word_bytes(n) * (size_t)N);
    if (off_edges + 8 * 1024 > kScratchBytes)
        return set_error(IGC_ERR_INVALID, "N=%u strings need %zu bytes of packed words; the no-allocation build has "
                                          "%zu bytes of scratch", N, off_edges, kScratchBytes);
    const uint64_t cap = (kScratchBytes - off_edges) / sizeof(uint64_t);
    uint32_t *misc = reinterpret_cast<uint32_t *>(scr);
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(misc);
    void *packed = scr + off_packed;
    uint64_t *edges = reinterpret_cast<uint64_t *>(scr + off_edges);
    int rc;
    HIP_TRY(hipEventRecord(ctx->ev[0], s));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, misc, (uint64_t)kMiscWords, 0u);
    if ((rc = launch_pack(s, d_syms, N, n, packed, misc + 2))) return rc;
    HIP_TRY(hipEventRecord(ctx->ev[1], s));
    if ((rc = launch_pairs(ctx_pred(ctx, n, K), n, K, sh.cpl, true, s, tiles, packed, G, 0u, edges, cap, counter)))
        return rc;
    HIP_TRY(hipEventRecord(ctx->ev[2], s));
    uint32_t host_misc[kMiscWords];
    HIP_TRY(hipMemcpyAsync(host_misc, misc, sizeof host_misc, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (host_misc[2]) return bad_symbols(host_misc[2]);
    const uint64_t E = (uint64_t)host_misc[0] | ((uint64_t)host_misc[1] << 32);
    *n_edges = E;
    if (2 * E > col_cap)
        return set_error(IGC_ERR_INVALID, "the graph has %llu edges: col needs %llu entries, the caller gave %llu",
                         (unsigned long long)E, (unsigned long long)(2 * E), (unsigned long long)col_cap);
    if (E == 0) {
        ctx->last = tm;
        return empty_csr();
    }
    // CSR by the segmented radix form, temporaries after the E edges
    uint32_t bsh = 1;
    while (bsh < 63 && ((uint64_t)1 << bsh) < N) ++bsh;
    const uint32_t rowbits = std::min(bsh, 32u - std::min(bsh, 31u));
    const uint32_t hb = bsh - rowbits;
    if (hb > kRadixMaxBits) return set_error(IGC_ERR_INVALID, "N=%u too large for the no-allocation build", N);
    SegPlan P;
    if ((rc = seg_plan(N, E, false, bsh, rowbits, hb, P))) return rc;
    size_t off = align256(off_edges + sizeof(uint64_t) * (size_t)E);
    auto carve = [&](size_t bytes) {
        uint8_t *p = scr + off;
        off = align256(off + bytes);
        return p;
    };
    SegBufs B{};
    B.row_ptr = d_row_ptr;
    B.col = d_col;
    B.keys0 = reinterpret_cast<uint32_t *>(carve(P.bytes_keys0));
    B.keys1 = reinterpret_cast<uint32_t *>(carve(P.bytes_keys1));
    B.rcount = reinterpret_cast<uint32_t *>(carve(P.bytes_rcount));
    B.roffset = reinterpret_cast<uint64_t *>(carve(P.bytes_roffset));
    B.bsum = reinterpret_cast<uint64_t *>(carve(P.bytes_bsum));
    B.segmeta = reinterpret_cast<SegMeta *>(carve(P.bytes_segmeta));
    if (off > kScratchBytes)
        return set_error(IGC_ERR_INVALID, "the graph's %llu edges need %zu bytes of scratch; the no-allocation build "
                                          "has %zu", (unsigned long long)E, off, kScratchBytes);
    uint32_t launches = 3;
    launch_csr_seg_into(s, P, B, edges, counter, misc, &launches, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[3], s));
    HIP_TRY(hipEventSynchronize(ctx->ev[3]));
    HIP_TRY(hipEventElapsedTime(&tm.pack_ms, ctx->ev[0], ctx->ev[1]));
    HIP_TRY(hipEventElapsedTime(&tm.pairs_ms, ctx->ev[1], ctx->ev[2]));
    HIP_TRY(hipEventElapsedTime(&tm.csr_ms, ctx->ev[2], ctx->ev[3]));
    HIP_TRY(hipEventElapsedTime(&tm.total_ms, ctx->ev[0], ctx->ev[3]));
    tm.launches = launches;
    tm.n_edges = E;
    ctx->last = tm;
    return IGC_OK;
}

} // namespace igc

using namespace igc;

extern "C" {

IGC_API int igc_graph_build_device_noalloc(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n,
                                           uint32_t edit_dist, uint64_t *d_row_ptr, uint32_t *d_col,
                                           uint64_t col_cap, uint64_t *n_edges)
{
    if (!ctx || !n_edges || !d_row_ptr || (!d_syms && N) || (!d_col && col_cap))
        return set_error(IGC_ERR_INVALID, "null argument");
    if (n < 1 || n > IGC_MAX_LEN) return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%d", n, IGC_MAX_LEN);
    if (N > 0x7FFFFFFFu) return set_error(IGC_ERR_INVALID, "N=%u too large", N);
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = g_scratch[ctx->device] ? IGC_OK : scratch_attach(ctx->device); // set up at context creation
    if (rc) return rc;
    uint8_t *scr = g_scratch[ctx->device];
    std::lock_guard<std::mutex> dk(g_device_mu[ctx->device]);
    return noalloc_locked(ctx, scr, d_syms, N, n, edit_dist, d_row_ptr, d_col, col_cap, n_edges);
}

} // extern "C"
