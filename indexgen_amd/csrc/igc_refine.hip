// igc_refine.hip -- cluster-based refinement of a candidate set, end to end on
// one device context: the reference's clustered codebook loop
// (GenerateCodebookAdj, src/SparseMat.cpp:918-1095) with its pieces replaced:
//   cluster step      FP32 K-means on one-hot features, Lloyd iterations
//                     replayed from a captured hipGraph (igc_kmeans.hip)
//   per-cluster solve all clusters' conflict graphs in one batched build (one
//                     launch per pair predicate, igc_batch.hip), then the greedy per cluster on host
//                     threads (the reference runs one Python process + GPU
//                     context per cluster, :985-1020 -> :784-851)
//   combine           next candidates = cluster codebooks concatenated in
//                     cluster order (:1029-1032)
//   convergence       stop when the last `converge_n` sizes are equal
//                     (:1063-1090), or after max_rounds
#include "igc_device.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

using namespace igc;

extern "C" {

IGC_API int igc_refine(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist, uint32_t k,
                       uint32_t kmeans_iters, uint32_t max_rounds, uint32_t converge_n, uint64_t seed,
                       uint32_t host_threads, uint32_t *out_idx, uint32_t *out_count, igc_refine_stats *stats)
{
    if (!ctx || (!syms && N) || !out_idx || !out_count) return set_error(IGC_ERR_INVALID, "null argument");
    if (n < 1 || n > IGC_MAX_QUATERNARY_LEN)
        return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%d (K-means runs on quaternary strings)", n,
                         IGC_MAX_QUATERNARY_LEN);
    if (k < 1) return set_error(IGC_ERR_INVALID, "k must be >= 1");
    if (converge_n < 1) converge_n = 1;
    if (host_threads < 1) host_threads = 1;
    igc_refine_stats st{};
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    int rc;
    if ((rc = kmeans_attach(ctx))) return rc;

    std::vector<uint32_t> cur(N);
    for (uint32_t i = 0; i < N; ++i) cur[i] = i;
    std::vector<uint32_t> sizes;
    std::vector<uint8_t> cur_syms;
    std::vector<int32_t> assign;
    using clk = std::chrono::steady_clock;

    for (uint32_t round = 0; round < max_rounds && !cur.empty(); ++round) {
        const uint32_t M = (uint32_t)cur.size();
        // the reference raises k so no cluster exceeds 200,000 strings on average
        uint32_t keff = k;
        if (M / keff > 200000u) keff = std::max<uint32_t>(2, (M + 199999u) / 200000u);
        keff = std::min(keff, M);
        cur_syms.resize((size_t)M * n);
        for (uint32_t i = 0; i < M; ++i) std::memcpy(&cur_syms[(size_t)i * n], syms + (size_t)cur[i] * n, n);

        // 1. cluster (device): K-means with graph-replayed Lloyd iterations
        const auto t0 = clk::now();
        if ((rc = ctx->syms.ensure(std::max<size_t>(cur_syms.size(), 1)))) return rc;
        HIP_TRY(hipMemcpyAsync(ctx->syms.p, cur_syms.data(), cur_syms.size(), hipMemcpyHostToDevice, ctx->stream));
        float km_ms = 0;
        if ((rc = kmeans_device_locked(ctx, ctx->syms.as<uint8_t>(), M, n, keff, kmeans_iters,
                                       seed + 0x9E37ull * round, true, &km_ms)))
            return rc;
        assign.resize(M);
        HIP_TRY(hipMemcpyAsync(assign.data(), kmeans_assign_ptr(ctx), sizeof(int32_t) * M, hipMemcpyDeviceToHost,
                               ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
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
        std::vector<uint8_t> grouped((size_t)M * n);
        for (uint32_t q = 0; q < M; ++q) std::memcpy(&grouped[(size_t)q * n], &cur_syms[(size_t)members[q] * n], n);
        std::vector<igc_set> sets(nc);
        for (uint32_t t = 0; t < nc; ++t) {
            const uint32_t c = cl[t];
            sets[t] = igc_set{grouped.data() + (size_t)cstart[c] * n, ccount[c], n, edit_dist, 0u};
        }

        // 2. all clusters' conflict graphs in one launch
        std::vector<igc_csr> g(nc);
        if ((rc = batch_build_host_locked(ctx, sets.data(), nc, g.data()))) return rc;
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
    std::memcpy(out_idx, cur.data(), sizeof(uint32_t) * cur.size());
    *out_count = (uint32_t)cur.size();
    if (stats) *stats = st;
    return IGC_OK;
}

} // extern "C"
