// igc_refine.hip -- cluster-based refinement of a candidate set, end to end on
// one device context: the reference's clustered codebook loop
// (GenerateCodebookAdj, src/SparseMat.cpp:918-1095) with its pieces replaced:
//   cluster step      FP32 K-means on one-hot features (igc_kmeans.hip)
//   per-cluster solve all clusters' conflict graphs in one batched build (one
//                     launch per pair predicate, igc_batch.hip), then the greedy
//                     per cluster on host threads (the reference runs one Python
//                     process + GPU context per cluster, :985-1020 -> :784-851)
//   combine           next candidates = cluster codebooks concatenated in
//                     cluster order (:1029-1032)
//   convergence       stop when the last `converge_n` sizes are equal
//                     (:1063-1090), or after max_rounds
//
// Device state across rounds.  The candidates' symbols go to the device once;
// a round's candidates are an index list into them (4 B per string up, the
// assignment 4 B per string down, then the cluster-ordered index list up).  A
// round's device work up to the assignment -- gather the round's strings,
// pack, seed the centroids, the Lloyd iterations, the final assignment -- is
// captured once as a hipGraph with grids for the first (largest) round and
// the round's string count read from device memory, and replayed in every
// round (rounds shrink, so the same graph serves them all; it is captured again
// only if the cluster count changes).  The batched build's launches follow the
// round's cluster sizes and the edge count read back before the CSR, so they
// are launched directly; they read the cluster-ordered strings from the device.
#include "igc_device.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

namespace igc {

// out[q] = strings[idx[q]] for q < *dN (n bytes each).
static __global__ void k_gather_rows(const uint8_t *__restrict__ strings, const uint32_t *__restrict__ idx,
                                     const uint32_t *__restrict__ dN, uint32_t n, uint8_t *__restrict__ out)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *dN) return;
    const uint8_t *src = strings + (size_t)idx[q] * n;
    uint8_t *dst = out + (size_t)q * n;
    for (uint32_t p = 0; p < n; ++p) dst[p] = src[p];
}

// Device buffers of one refinement run (charged to the context's budget
// slot), freed once the stream has finished with them.
struct RefineBufs {
    DevBuf all, idx, cur, grouped, count;
    hipStream_t stream;
    explicit RefineBufs(igc_ctx *ctx) : stream(ctx->stream)
    {
        for (DevBuf *b : {&all, &idx, &cur, &grouped, &count}) b->slot = ctx->mem.pool ? &ctx->mem : nullptr;
    }
    ~RefineBufs()
    {
        (void)hipStreamSynchronize(stream);
        for (DevBuf *b : {&all, &idx, &cur, &grouped, &count}) b->release();
    }
};

struct RoundGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    uint32_t k = 0;
    void reset()
    {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        ge = nullptr;
        g = nullptr;
        k = 0;
    }
    ~RoundGraph() { reset(); }
};

// The refinement loop on a locked context (syms: host memory, N x n).
int refine_locked(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist, uint32_t k,
                  uint32_t kmeans_iters, uint32_t max_rounds, uint32_t converge_n, uint64_t seed,
                  uint32_t host_threads, bool use_graph, uint32_t *out_idx, uint32_t *out_count,
                  igc_refine_stats *stats)
{
    if (n < 1 || n > IGC_MAX_QUATERNARY_LEN)
        return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%d (K-means runs on quaternary strings)", n,
                         IGC_MAX_QUATERNARY_LEN);
    if (k < 1) return set_error(IGC_ERR_INVALID, "k must be >= 1");
    if (converge_n < 1) converge_n = 1;
    if (host_threads < 1) host_threads = 1;
    igc_refine_stats st{};
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = kmeans_attach(ctx))) return rc;
    RefineBufs B(ctx);
    RoundGraph graph;
    const size_t bytes = (size_t)N * n;
    if ((rc = B.all.ensure(std::max<size_t>(bytes, 1)))) return rc;
    if ((rc = B.idx.ensure(sizeof(uint32_t) * std::max<size_t>(N, 1)))) return rc;
    if ((rc = B.cur.ensure(std::max<size_t>(bytes, 1)))) return rc;
    if ((rc = B.grouped.ensure(std::max<size_t>(bytes, 1)))) return rc;
    if ((rc = B.count.ensure(sizeof(uint32_t) * 4))) return rc;
    if (bytes) HIP_TRY(hipMemcpyAsync(B.all.p, syms, bytes, hipMemcpyHostToDevice, s));
    uint32_t *d_count = B.count.as<uint32_t>();

    std::vector<uint32_t> cur(N);
    for (uint32_t i = 0; i < N; ++i) cur[i] = i;
    std::vector<uint32_t> sizes, seeds, gidx;
    std::vector<int32_t> assign;
    uint32_t round_n = 0; // the round's string count, copied to d_count
    using clk = std::chrono::steady_clock;

    for (uint32_t round = 0; round < max_rounds && !cur.empty(); ++round) {
        const uint32_t M = (uint32_t)cur.size();
        // the reference raises k so no cluster exceeds 200,000 strings on average
        uint32_t keff = k;
        if (M / keff > 200000u) keff = std::max<uint32_t>(2, (M + 199999u) / 200000u);
        keff = std::min(keff, M);

        // 1. cluster (device): the captured K-means, replayed
        const auto t0 = clk::now();
        if ((rc = kmeans_prepare_locked(ctx, N, n, keff))) return rc; // sized for the first round: stays in place
        round_n = M;
        HIP_TRY(hipEventRecord(ctx->ev[0], s));
        HIP_TRY(hipMemcpyAsync(B.idx.p, cur.data(), sizeof(uint32_t) * M, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_count, &round_n, sizeof(uint32_t), hipMemcpyHostToDevice, s));
        if ((rc = kmeans_seed_locked(ctx, M, keff, seed + 0x9E37ull * round, seeds))) return rc;
        auto launch_round = [&](uint32_t grid_n) {
            hipLaunchKernelGGL(k_gather_rows, dim3((grid_n + 255) / 256), dim3(256), 0, s, B.all.as<uint8_t>(),
                               B.idx.as<uint32_t>(), d_count, n, B.cur.as<uint8_t>());
            kmeans_launch_locked(ctx, B.cur.as<uint8_t>(), grid_n, d_count, n, keff, kmeans_iters);
        };
        if (use_graph) {
            if (!graph.ge || graph.k != keff) {
                graph.reset();
                HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                launch_round(N);
                hipError_t e = hipStreamEndCapture(s, &graph.g);
                if (e != hipSuccess) return set_error(IGC_ERR_DEVICE, "round capture failed: %s", hipGetErrorString(e));
                e = hipGraphInstantiate(&graph.ge, graph.g, nullptr, nullptr, 0);
                if (e != hipSuccess) return set_error(IGC_ERR_DEVICE, "round instantiate failed: %s", hipGetErrorString(e));
                graph.k = keff;
            }
            HIP_TRY(hipGraphLaunch(graph.ge, s));
        } else {
            launch_round(M);
            HIP_TRY(hipGetLastError());
        }
        HIP_TRY(hipEventRecord(ctx->ev[1], s));
        uint32_t flags[kMiscWords];
        assign.resize(M);
        HIP_TRY(hipMemcpyAsync(flags, kmeans_flags_ptr(ctx), sizeof flags, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(assign.data(), kmeans_assign_ptr(ctx), sizeof(int32_t) * M, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (flags[2]) return set_error(IGC_ERR_INVALID, "symbol code above 3 in input");
        float km_ms = 0;
        HIP_TRY(hipEventElapsedTime(&km_ms, ctx->ev[0], ctx->ev[1]));
        const auto t1 = clk::now();

        // clusters in centroid order, members in candidate order; empty ones dropped
        std::vector<uint32_t> ccount(keff, 0);
        for (uint32_t i = 0; i < M; ++i) ccount[(uint32_t)assign[i]]++;
        std::vector<uint32_t> cstart(keff + 1, 0);
        for (uint32_t c = 0; c < keff; ++c) cstart[c + 1] = cstart[c] + ccount[c];
        std::vector<uint32_t> members(M), fill(cstart.begin(), cstart.end() - 1);
        for (uint32_t i = 0; i < M; ++i) members[fill[(uint32_t)assign[i]]++] = i;
        std::vector<uint32_t> cl;
        for (uint32_t c = 0; c < keff; ++c)
            if (ccount[c]) cl.push_back(c);
        const uint32_t nc = (uint32_t)cl.size();
        // the strings in cluster order, gathered on the device from the resident set
        gidx.resize(M);
        for (uint32_t q = 0; q < M; ++q) gidx[q] = cur[members[q]];
        HIP_TRY(hipMemcpyAsync(B.idx.p, gidx.data(), sizeof(uint32_t) * M, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gather_rows, dim3((M + 255) / 256), dim3(256), 0, s, B.all.as<uint8_t>(),
                           B.idx.as<uint32_t>(), d_count, n, B.grouped.as<uint8_t>());
        HIP_TRY(hipGetLastError());
        std::vector<igc_set> sets(nc);
        for (uint32_t t = 0; t < nc; ++t) {
            const uint32_t c = cl[t];
            sets[t] = igc_set{B.grouped.as<uint8_t>() + (size_t)cstart[c] * n, ccount[c], n, edit_dist, 0u};
        }

        // 2. all clusters' conflict graphs in one batched build
        std::vector<igc_csr> g(nc);
        if ((rc = batch_build_dev_syms_locked(ctx, B.grouped.as<uint8_t>(), sets.data(), nc, g.data()))) return rc;
        const float pairs_ms = ctx->last.pairs_ms;
        const auto t2 = clk::now();

        // 3. greedy per cluster on host threads (clusters are independent)
        std::vector<std::vector<uint32_t>> picks(nc);
        std::vector<int> errs(host_threads, IGC_OK);
        std::vector<std::thread> pool;
        for (uint32_t w = 0; w < host_threads; ++w) {
            pool.emplace_back([&, w] {
                for (uint32_t t = w; t < nc; t += host_threads) {
                    picks[t].resize(g[t].n_nodes);
                    uint32_t cnt = 0;
                    const int e = igc_greedy(&g[t], picks[t].data(), &cnt);
                    if (e) errs[w] = e;
                    picks[t].resize(cnt);
                }
            });
        }
        for (auto &th : pool) th.join();
        for (auto &x : g) igc_csr_free(&x);
        for (int e : errs)
            if (e) return e;
        const auto t3 = clk::now();

        // 4. combine in cluster order
        std::vector<uint32_t> next;
        for (uint32_t t = 0; t < nc; ++t) {
            const uint32_t base = cstart[cl[t]];
            for (uint32_t v : picks[t]) next.push_back(cur[members[base + v]]);
        }
        sizes.push_back((uint32_t)next.size());
        if (st.rounds < IGC_REFINE_MAX_ROUNDS) {
            igc_refine_round &r = st.round[st.rounds];
            r.candidates = M;
            r.clusters = nc;
            r.codewords = (uint32_t)next.size();
            r.kmeans_ms = km_ms;
            r.pairs_ms = pairs_ms;
            r.cluster_s = std::chrono::duration<double>(t1 - t0).count();
            r.graph_s = std::chrono::duration<double>(t2 - t1).count();
            r.greedy_s = std::chrono::duration<double>(t3 - t2).count();
            uint64_t pairs = 0;
            for (uint32_t t = 0; t < nc; ++t) pairs += (uint64_t)sets[t].N * (sets[t].N - 1) / 2;
            r.pairs = pairs;
        }
        st.rounds++;
        cur.swap(next);
        bool converged = sizes.size() >= converge_n;
        for (uint32_t q = 1; converged && q < converge_n; ++q)
            if (sizes[sizes.size() - 1] != sizes[sizes.size() - 1 - q]) converged = false;
        if (converged) {
            st.converged = 1;
            break;
        }
    }
    HIP_TRY(hipStreamSynchronize(s));
    std::memcpy(out_idx, cur.data(), sizeof(uint32_t) * cur.size());
    *out_count = (uint32_t)cur.size();
    if (stats) *stats = st;
    return IGC_OK;
}

} // namespace igc

using namespace igc;

extern "C" {

IGC_API int igc_refine(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist, uint32_t k,
                       uint32_t kmeans_iters, uint32_t max_rounds, uint32_t converge_n, uint64_t seed,
                       uint32_t host_threads, uint32_t *out_idx, uint32_t *out_count, igc_refine_stats *stats)
{
    if (!ctx || (!syms && N) || !out_idx || !out_count) return set_error(IGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return refine_locked(ctx, syms, N, n, edit_dist, k, kmeans_iters, max_rounds, converge_n, seed, host_threads,
                         ctx->graph_replay, out_idx, out_count, stats);
}

} // extern "C"
