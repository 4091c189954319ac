// igc_kmeans.hip -- FP32 K-means on one-hot symbol features, for cluster-based
// refinement of a candidate set.
//
// Feature of a length-n string: the one-hot 4n-vector (dimension 4p+s is 1 iff
// symbol p equals s).  For a centroid c, ||x - c||^2 = n - 2 sum_p C[c][4p+x_p]
// + ||c||^2, so the assignment needs n table lookups per centroid, not a 4n-wide
// dot product; the table (k x 4n floats) is staged through LDS in chunks.
// Determinism: every sum runs in a fixed order with FP contraction off, ties go
// to the lowest centroid index, and the update uses exact integer histograms
// (cnt[c][4p+s] / size[c], one correctly rounded division), so the result does
// not depend on scheduling or on which other work shares the GPU.
//
// Replaces the reference's random-centre assignment (RandomCluster::fit,
// src/clustering/impl/random_cluster/RandomCluster.cpp:42-136, whose distances
// come from src/gpu_cluster.py:16-193,287) with Lloyd iterations captured in a
// hipGraph and replayed.
#include "igc_device.hpp"

#include <cstring>
#include <vector>

#pragma clang fp contract(off)

namespace igc {

constexpr int kKmThreads = 256;

// Symbol p of a packed word (32-bit words for n <= 16, 64-bit above).
template <class W>
__device__ __forceinline__ uint32_t sym_at(W w, uint32_t p)
{
    constexpr int H = WordTraits<W>::half;
    return (uint32_t)((w >> p) & 1u) | ((uint32_t)((w >> (H + p)) & 1u) << 1);
}

// assign[i] = argmin_c (cn2[c] - 2 * sum_p C[c][4p + s_p]), first index on ties.
// The centroid table goes through LDS in chunks of kc centroids, transposed:
// row j = 4p + s holds C[c][j] for the chunk's centroids side by side (rows
// km_stride(kc) floats apart, 4 more than the chunk rounded up to 4, so the four
// rows one position can select start in different banks), and one 16-byte read
// gives a string's term for four centroids.  Padding centroids have norm
// +infinity and never win.  dN (may be null) holds the point count in device
// memory instead of N, so a launch captured for the largest round serves every
// round (grid-stride loops).
__host__ __device__ inline uint32_t km_stride(uint32_t kc) { return ((kc + 3u) & ~3u) + 4u; }

template <class W>
__global__ __launch_bounds__(kKmThreads) void k_km_assign(const W *__restrict__ packed, uint32_t N,
                                                                  const uint32_t *__restrict__ dN, uint32_t n,
                                                                  const float *__restrict__ C,
                                                                  const float *__restrict__ cn2, uint32_t k,
                                                                  uint32_t kc, int32_t *__restrict__ assign)
{
    extern __shared__ float4 sh4[];
    float *sh = reinterpret_cast<float *>(sh4);
    if (dN) N = *dN;
    const uint32_t D = 4 * n;
    const uint32_t S = km_stride(kc);
    float *sT = sh;
    float *sN = sh + (size_t)D * S;
    for (uint32_t base = blockIdx.x * blockDim.x; base < N; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const W w = i < N ? packed[i] : (W)0;
        constexpr uint32_t P = WordTraits<W>::half;
        uint32_t off[P];
#pragma unroll
        for (uint32_t p = 0; p < P; ++p) off[p] = (4 * p + sym_at(w, p)) * S;
        float best = 0.0f;
        int32_t bi = -1;
        for (uint32_t c0 = 0; c0 < k; c0 += kc) {
            const uint32_t m = min(kc, k - c0);
            const uint32_t m4 = (m + 3u) & ~3u;
            __syncthreads();
            for (uint32_t t = threadIdx.x; t < m4 * D; t += blockDim.x) {
                const uint32_t c = t / D, j = t - c * D;
                sT[(size_t)j * S + c] = c < m ? C[(size_t)(c0 + c) * D + j] : 0.0f;
            }
            for (uint32_t t = threadIdx.x; t < m4; t += blockDim.x) sN[t] = t < m ? cn2[c0 + t] : INFINITY;
            __syncthreads();
            // four centroids per read, the symbol loop unrolled to the word's
            // length; each sum keeps its order, p ascending
            for (uint32_t c = 0; c < m4; c += 4) {
                float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll
                for (uint32_t p = 0; p < P; ++p) {
                    if (p < n) {
                        const float4 v = *reinterpret_cast<const float4 *>(sT + off[p] + c);
                        a0 = a0 + v.x;
                        a1 = a1 + v.y;
                        a2 = a2 + v.z;
                        a3 = a3 + v.w;
                    }
                }
                const float a[4] = {a0, a1, a2, a3};
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) {
                    const float twice = 2.0f * a[q];
                    const float dist = sN[c + q] - twice;
                    if (bi < 0 || dist < best) {
                        best = dist;
                        bi = (int32_t)(c0 + c + q);
                    }
                }
            }
        }
        if (i < N) assign[i] = bi;
    }
}

// LDS bytes of k_km_assign for D = 4n dimensions and chunks of kc centroids.
static size_t km_assign_lds(uint32_t D, uint32_t kc)
{
    return sizeof(float) * ((size_t)D * km_stride(kc) + ((kc + 3u) & ~3u));
}

// Integer histograms per cluster: cnt[c][4p+s] and size[c].  Each workgroup
// counts in LDS with two 16-bit counters per word (a workgroup sees at most
// 32768 points, so no half overflows) and flushes nonzero entries with one
// global atomic each.
template <class W>
__global__ __launch_bounds__(kKmThreads) void k_km_accum(const W *__restrict__ packed, uint32_t N,
                                                                 const uint32_t *__restrict__ dN, uint32_t n,
                                                                 const int32_t *__restrict__ assign,
                                                                 uint32_t k, uint32_t *__restrict__ cnt,
                                                                 uint32_t *__restrict__ size)
{
    extern __shared__ uint32_t hs[];
    if (dN) N = *dN;
    const uint32_t D = 4 * n;
    const uint32_t words = (k * D + 1) / 2;
    uint32_t *hc = hs;
    uint32_t *hz = hs + words;
    for (uint32_t t = threadIdx.x; t < words + k; t += blockDim.x) hs[t] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const W w = packed[i];
        const uint32_t c = (uint32_t)assign[i];
        for (uint32_t p = 0; p < n; ++p) {
            const uint32_t t = c * D + 4 * p + sym_at(w, p);
            atomicAdd(&hc[t >> 1], 1u << (16 * (t & 1)));
        }
        atomicAdd(&hz[c], 1u);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < k * D; t += blockDim.x) {
        const uint32_t v = (hc[t >> 1] >> (16 * (t & 1))) & 0xFFFFu;
        if (v) atomicAdd(&cnt[t], v);
    }
    for (uint32_t t = threadIdx.x; t < k; t += blockDim.x)
        if (hz[t]) atomicAdd(&size[t], hz[t]);
}

// C[c][d] = cnt[c][d] / size[c] (an empty cluster keeps its centroid), then
// cn2[c] = sum_d C[c][d]^2 in dimension order.  One workgroup per centroid.
static __global__ void k_km_update(const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ size, uint32_t D,
                                   float *__restrict__ C, float *__restrict__ cn2)
{
    const uint32_t c = blockIdx.x;
    const uint32_t z = size[c];
    for (uint32_t d = threadIdx.x; d < D; d += blockDim.x)
        if (z) C[(size_t)c * D + d] = (float)cnt[(size_t)c * D + d] / (float)z;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (uint32_t d = 0; d < D; ++d) {
            const float v = C[(size_t)c * D + d];
            const float sq = v * v;
            s = s + sq;
        }
        cn2[c] = s;
    }
}

template <class W>
__global__ void k_km_init(const W *__restrict__ packed, const uint32_t *__restrict__ seeds, uint32_t n, uint32_t k,
                          float *__restrict__ C, float *__restrict__ cn2)
{
    const uint32_t c = blockIdx.x;
    const uint32_t D = 4 * n;
    const W w = packed[seeds[c]];
    for (uint32_t d = threadIdx.x; d < D; d += blockDim.x) C[(size_t)c * D + d] = (sym_at(w, d >> 2) == (d & 3)) ? 1.0f : 0.0f;
    if (threadIdx.x == 0) cn2[c] = (float)n;
}

// cn2[c] = sum_d C[c][d]^2 in dimension order (as k_km_update computes it), for
// caller-provided centroids.  One thread per centroid.
static __global__ void k_km_norms(const float *__restrict__ C, uint32_t k, uint32_t D, float *__restrict__ cn2)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    float s = 0.0f;
    for (uint32_t d = 0; d < D; ++d) {
        const float v = C[(size_t)c * D + d];
        const float sq = v * v;
        s = s + sq;
    }
    cn2[c] = s;
}

// k distinct seed indices: a partial Fisher-Yates shuffle of 0..N-1 driven by
// SplitMix64(seed); seed index c is the c-th drawn.
void kmeans_seeds(uint32_t N, uint32_t k, uint64_t seed, std::vector<uint32_t> &out)
{
    std::vector<uint32_t> idx(N);
    for (uint32_t i = 0; i < N; ++i) idx[i] = i;
    uint64_t st = seed;
    out.resize(k);
    for (uint32_t i = 0; i < k; ++i) {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const uint32_t j = i + (uint32_t)(z % (uint64_t)(N - i));
        std::swap(idx[i], idx[j]);
        out[i] = idx[i];
    }
}

// K-means buffers: their own packed words and flags (not the context's, which
// the graph builds regrow), so launches captured with them stay valid.
struct KmBufs {
    DevBuf C, cn2, cnt, size, seeds, assign, packed, flags;
};

static KmBufs &km_bufs(igc_ctx *ctx) { return *reinterpret_cast<KmBufs *>(ctx->km); }

static uint32_t km_chunk(uint32_t D) { return std::max<uint32_t>(1, std::min<uint32_t>(256, 8192 / D)); }

static uint32_t km_grid(igc_ctx *ctx, uint32_t N)
{
    // at least ceil(N / 32768) workgroups so each sees <= 32768 points in k_km_accum
    return std::max<uint32_t>(grid_for(N, kKmThreads, (uint32_t)ctx->cu_count * 4), (N + 32767) / 32768);
}

// Packs the strings for K-means (the count in dN when non-null, else N).
template <class W>
__global__ void k_km_pack(const uint8_t *__restrict__ syms, uint32_t N, const uint32_t *__restrict__ dN, uint32_t n,
                          W *__restrict__ packed, uint32_t *__restrict__ bad)
{
    if (dN) N = *dN;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint8_t *p = syms + (size_t)i * n;
    uint32_t or_all = 0;
    for (uint32_t q = 0; q < n; ++q) or_all |= p[q];
    if constexpr (sizeof(W) == 8) packed[i] = pack_word64(p, n);
    else packed[i] = pack_word(p, n);
    if (or_all > 3u) atomicOr(bad, 1u);
}

// Launches one Lloyd iteration on ctx->stream (capturable: kernels only).
// Grids are sized for N points; dN (may be null) holds the actual count.
template <class W>
static void lloyd_iteration(igc_ctx *ctx, KmBufs &b, uint32_t N, const uint32_t *dN, uint32_t n, uint32_t k)
{
    hipStream_t s = ctx->stream;
    const uint32_t D = 4 * n;
    const uint32_t kc = km_chunk(D);
    const size_t lds_assign = km_assign_lds(D, kc);
    const size_t lds_accum = sizeof(uint32_t) * (((size_t)k * D + 1) / 2 + k);
    const uint32_t grid = km_grid(ctx, N);
    hipLaunchKernelGGL(k_km_assign<W>, dim3(grid), dim3(kKmThreads), lds_assign, s, b.packed.as<W>(), N, dN, n,
                       b.C.as<float>(), b.cn2.as<float>(), k, kc, b.assign.as<int32_t>());
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((uint64_t)k * D, 256, 1024)), dim3(256), 0, s, b.cnt.as<uint32_t>(),
                       (uint64_t)k * D, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, b.size.as<uint32_t>(), (uint64_t)k, 0u);
    hipLaunchKernelGGL(k_km_accum<W>, dim3(grid), dim3(kKmThreads), lds_accum, s, b.packed.as<W>(), N, dN, n,
                       b.assign.as<int32_t>(), k, b.cnt.as<uint32_t>(), b.size.as<uint32_t>());
    hipLaunchKernelGGL(k_km_update, dim3(k), dim3(64), 0, s, b.cnt.as<uint32_t>(), b.size.as<uint32_t>(), D,
                       b.C.as<float>(), b.cn2.as<float>());
}

static int km_check(uint32_t N, uint32_t n, uint32_t k)
{
    if (k < 1 || k > N) return set_error(IGC_ERR_INVALID, "k=%u must be in 1..N (N=%u)", k, N);
    if (((size_t)k * 4 * n + 1) / 2 * 4 + (size_t)k * 4 > 64 * 1024)
        return set_error(IGC_ERR_INVALID, "k=%u too large for the LDS histogram (k*4n/2 + k words <= 16384)", k);
    return IGC_OK;
}

// Device buffers of a K-means of up to N points (they stay put afterwards, so
// launches captured with them can be replayed).
template <class W>
static int km_buffers(igc_ctx *ctx, KmBufs &b, uint32_t N, uint32_t n, uint32_t k)
{
    const uint32_t D = 4 * n;
    int rc;
    if ((rc = b.packed.ensure(sizeof(W) * (size_t)N))) return rc;
    if ((rc = b.flags.ensure(sizeof(uint32_t) * kMiscWords))) return rc;
    if ((rc = b.C.ensure(sizeof(float) * (size_t)k * D))) return rc;
    if ((rc = b.cn2.ensure(sizeof(float) * k))) return rc;
    if ((rc = b.cnt.ensure(sizeof(uint32_t) * (size_t)k * D))) return rc;
    if ((rc = b.size.ensure(sizeof(uint32_t) * k))) return rc;
    if ((rc = b.seeds.ensure(sizeof(uint32_t) * k))) return rc;
    if ((rc = b.assign.ensure(sizeof(int32_t) * (size_t)N))) return rc;
    return IGC_OK;
}

// The whole K-means as kernels only (capturable): pack, seeds -> centroids
// (b.seeds already filled), `iters` Lloyd iterations, then (with final) the
// assignment against the last centroids.  Grids for N points; dN (may be
// null) holds the actual count.
template <class W>
static void km_launch(igc_ctx *ctx, KmBufs &b, const uint8_t *d_syms, uint32_t N, const uint32_t *dN, uint32_t n,
                      uint32_t k, uint32_t iters, bool final = true)
{
    hipStream_t s = ctx->stream;
    const uint32_t D = 4 * n;
    uint32_t *misc = b.flags.as<uint32_t>();
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, misc, (uint64_t)kMiscWords, 0u);
    hipLaunchKernelGGL(k_km_pack<W>, dim3((N + 255) / 256), dim3(256), 0, s, d_syms, N, dN, n, b.packed.as<W>(),
                       misc + 2);
    hipLaunchKernelGGL(k_km_init<W>, dim3(k), dim3(64), 0, s, b.packed.as<W>(), b.seeds.as<uint32_t>(), n, k,
                       b.C.as<float>(), b.cn2.as<float>());
    for (uint32_t t = 0; t < iters; ++t) lloyd_iteration<W>(ctx, b, N, dN, n, k);
    if (!final) return;
    const uint32_t kc = km_chunk(D);
    hipLaunchKernelGGL(k_km_assign<W>, dim3(km_grid(ctx, N)), dim3(kKmThreads), km_assign_lds(D, kc),
                       s, b.packed.as<W>(), N, dN, n, b.C.as<float>(), b.cn2.as<float>(), k, kc,
                       b.assign.as<int32_t>());
}

// K-means on symbols already in device memory: seeds -> `iters` Lloyd
// iterations (assign, histogram, update) -> a final assign against the last
// centroids, left in the context's assign buffer.  With use_graph one Lloyd
// iteration is captured and replayed `iters` times.
template <class W>
static int kmeans_device_w(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n, uint32_t k, uint32_t iters,
                           uint64_t seed, bool use_graph, float *ms_out)
{
    int rc;
    if ((rc = km_check(N, n, k))) return rc;
    KmBufs &b = km_bufs(ctx);
    if ((rc = km_buffers<W>(ctx, b, N, n, k))) return rc;
    std::vector<uint32_t> seeds;
    kmeans_seeds(N, k, seed, seeds);
    hipStream_t s = ctx->stream;
    uint32_t *misc = b.flags.as<uint32_t>();
    HIP_TRY(hipEventRecord(ctx->ev[0], s));
    HIP_TRY(hipMemcpyAsync(b.seeds.p, seeds.data(), sizeof(uint32_t) * k, hipMemcpyHostToDevice, s));
    if (!use_graph || iters == 0) {
        km_launch<W>(ctx, b, d_syms, N, nullptr, n, k, iters);
        HIP_TRY(hipGetLastError());
    } else {
        km_launch<W>(ctx, b, d_syms, N, nullptr, n, k, 0, false); // pack, seeds
        HIP_TRY(hipGetLastError());
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        lloyd_iteration<W>(ctx, b, N, nullptr, n, k);
        hipError_t e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) return set_error(IGC_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return set_error(IGC_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
        }
        for (uint32_t t = 0; t < iters && e == hipSuccess; ++t) e = hipGraphLaunch(ge, s);
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return set_error(IGC_ERR_DEVICE, "graph launch failed: %s", hipGetErrorString(e));
        // final assignment against the last centroids
        const uint32_t D = 4 * n, kc = km_chunk(D);
        hipLaunchKernelGGL(k_km_assign<W>, dim3(km_grid(ctx, N)), dim3(kKmThreads), km_assign_lds(D, kc), s,
                           b.packed.as<W>(), N, nullptr, n,
                           b.C.as<float>(), b.cn2.as<float>(), k, kc, b.assign.as<int32_t>());
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev[1], s));
    uint32_t host_misc[kMiscWords];
    HIP_TRY(hipMemcpyAsync(host_misc, misc, sizeof host_misc, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (host_misc[2]) return set_error(IGC_ERR_INVALID, "symbol code above 3 in input");
    if (ms_out) HIP_TRY(hipEventElapsedTime(ms_out, ctx->ev[0], ctx->ev[1]));
    return IGC_OK;
}

int kmeans_device_locked(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n, uint32_t k, uint32_t iters,
                         uint64_t seed, bool use_graph, float *ms_out)
{
    if (n < 1 || n > IGC_MAX_QUATERNARY_LEN)
        return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%d (K-means runs on quaternary strings)", n,
                         IGC_MAX_QUATERNARY_LEN);
    if (wide_words(n)) return kmeans_device_w<uint64_t>(ctx, d_syms, N, n, k, iters, seed, use_graph, ms_out);
    return kmeans_device_w<uint32_t>(ctx, d_syms, N, n, k, iters, seed, use_graph, ms_out);
}

int kmeans_prepare_locked(igc_ctx *ctx, uint32_t Nmax, uint32_t n, uint32_t k)
{
    int rc;
    if ((rc = kmeans_attach(ctx))) return rc;
    if ((rc = km_check(Nmax, n, k))) return rc;
    KmBufs &b = km_bufs(ctx);
    return wide_words(n) ? km_buffers<uint64_t>(ctx, b, Nmax, n, k) : km_buffers<uint32_t>(ctx, b, Nmax, n, k);
}

int kmeans_seed_locked(igc_ctx *ctx, uint32_t N, uint32_t k, uint64_t seed, std::vector<uint32_t> &host_seeds)
{
    kmeans_seeds(N, k, seed, host_seeds);
    HIP_TRY(hipMemcpyAsync(km_bufs(ctx).seeds.p, host_seeds.data(), sizeof(uint32_t) * k, hipMemcpyHostToDevice,
                           ctx->stream));
    return IGC_OK;
}

void kmeans_launch_locked(igc_ctx *ctx, const uint8_t *d_syms, uint32_t Nmax, const uint32_t *dN, uint32_t n,
                          uint32_t k, uint32_t iters)
{
    KmBufs &b = km_bufs(ctx);
    if (wide_words(n)) km_launch<uint64_t>(ctx, b, d_syms, Nmax, dN, n, k, iters);
    else km_launch<uint32_t>(ctx, b, d_syms, Nmax, dN, n, k, iters);
}

void kmeans_release(igc_ctx *ctx)
{
    if (!ctx->km) return;
    KmBufs *b = reinterpret_cast<KmBufs *>(ctx->km);
    for (DevBuf *d : {&b->C, &b->cn2, &b->cnt, &b->size, &b->seeds, &b->assign, &b->packed, &b->flags})
        d->release();
    delete b;
    ctx->km = nullptr;
}

int kmeans_attach(igc_ctx *ctx)
{
    if (!ctx->km) {
        KmBufs *b = new (std::nothrow) KmBufs();
        if (!b) return set_error(IGC_ERR_OOM, "host allocation failed");
        if (ctx->mem.pool)
            for (DevBuf *d : {&b->C, &b->cn2, &b->cnt, &b->size, &b->seeds, &b->assign, &b->packed, &b->flags})
                d->slot = &ctx->mem;
        ctx->km = b;
    }
    return IGC_OK;
}

int32_t *kmeans_assign_ptr(igc_ctx *ctx) { return reinterpret_cast<KmBufs *>(ctx->km)->assign.as<int32_t>(); }
uint32_t *kmeans_flags_ptr(igc_ctx *ctx) { return reinterpret_cast<KmBufs *>(ctx->km)->flags.as<uint32_t>(); }

} // namespace igc

using namespace igc;

extern "C" {

IGC_API int igc_kmeans(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t k, uint32_t iters,
                       uint64_t seed, uint32_t flags, int32_t *assign_out, float *centroids_out)
{
    if (!ctx || (!syms && N) || !assign_out) return set_error(IGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    int rc;
    if ((rc = kmeans_attach(ctx))) return rc;
    const size_t bytes = (size_t)N * n;
    if ((rc = ctx->syms.ensure(std::max<size_t>(bytes, 1)))) return rc;
    if (bytes) HIP_TRY(hipMemcpyAsync(ctx->syms.p, syms, bytes, hipMemcpyHostToDevice, ctx->stream));
    float ms = 0;
    if ((rc = kmeans_device_locked(ctx, ctx->syms.as<uint8_t>(), N, n, k, iters, seed, !(flags & IGC_KMEANS_EAGER),
                                   &ms)))
        return rc;
    KmBufs &b = *reinterpret_cast<KmBufs *>(ctx->km);
    HIP_TRY(hipMemcpyAsync(assign_out, b.assign.p, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToHost, ctx->stream));
    if (centroids_out)
        HIP_TRY(hipMemcpyAsync(centroids_out, b.C.p, sizeof(float) * (size_t)k * 4 * n, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->last = igc_timing{};
    ctx->last.total_ms = ms;
    return IGC_OK;
}

IGC_API int igc_assign_fp32(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, const float *centroids,
                            uint32_t k, int32_t *out)
{
    if (!ctx || (!syms && N) || !centroids || (!out && N)) return set_error(IGC_ERR_INVALID, "null argument");
    if (n < 1 || n > IGC_MAX_QUATERNARY_LEN)
        return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%d (K-means runs on quaternary strings)", n,
                         IGC_MAX_QUATERNARY_LEN);
    if (k < 1) return set_error(IGC_ERR_INVALID, "k must be at least 1");
    if (N == 0) return IGC_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    int rc;
    if ((rc = kmeans_attach(ctx))) return rc;
    KmBufs &b = km_bufs(ctx);
    hipStream_t s = ctx->stream;
    const uint32_t D = 4 * n;
    const size_t bytes = (size_t)N * n;
    const bool wide = wide_words(n);
    if ((rc = ctx->syms.ensure(bytes))) return rc;
    if ((rc = ctx->packed.ensure(word_bytes(n) * ((size_t)N + 1)))) return rc;
    if ((rc = ctx->misc.ensure(sizeof(uint32_t) * kMiscWords))) return rc;
    if ((rc = b.C.ensure(sizeof(float) * (size_t)k * D))) return rc;
    if ((rc = b.cn2.ensure(sizeof(float) * k))) return rc;
    if ((rc = b.assign.ensure(sizeof(int32_t) * (size_t)N))) return rc;
    uint32_t *misc = ctx->misc.as<uint32_t>();
    HIP_TRY(hipMemcpyAsync(ctx->syms.p, syms, bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b.C.p, centroids, sizeof(float) * (size_t)k * D, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, misc, (uint64_t)kMiscWords, 0u);
    if ((rc = launch_pack(s, ctx->syms.as<uint8_t>(), N, n, ctx->packed.p, misc + 2))) return rc;
    hipLaunchKernelGGL(k_km_norms, dim3((k + 255) / 256), dim3(256), 0, s, b.C.as<float>(), k, D, b.cn2.as<float>());
    const uint32_t kc = km_chunk(D);
    const size_t lds = km_assign_lds(D, kc);
    if (wide)
        hipLaunchKernelGGL(k_km_assign<uint64_t>, dim3(km_grid(ctx, N)), dim3(kKmThreads), lds, s,
                           ctx->packed.as<uint64_t>(), N, nullptr, n, b.C.as<float>(), b.cn2.as<float>(), k, kc,
                           b.assign.as<int32_t>());
    else
        hipLaunchKernelGGL(k_km_assign<uint32_t>, dim3(km_grid(ctx, N)), dim3(kKmThreads), lds, s,
                           ctx->packed.as<uint32_t>(), N, nullptr, n, b.C.as<float>(), b.cn2.as<float>(), k, kc,
                           b.assign.as<int32_t>());
    HIP_TRY(hipGetLastError());
    uint32_t host_misc[kMiscWords];
    HIP_TRY(hipMemcpyAsync(host_misc, misc, sizeof host_misc, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(out, b.assign.p, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (host_misc[2]) return set_error(IGC_ERR_INVALID, "symbol code above 3 in input");
    return IGC_OK;
}

IGC_API int igc_kmeans_seeds(uint32_t N, uint32_t k, uint64_t seed, uint32_t *out)
{
    if (!out || k > N) return set_error(IGC_ERR_INVALID, "bad arguments");
    std::vector<uint32_t> v;
    kmeans_seeds(N, k, seed, v);
    std::memcpy(out, v.data(), sizeof(uint32_t) * k);
    return IGC_OK;
}

} // extern "C"
