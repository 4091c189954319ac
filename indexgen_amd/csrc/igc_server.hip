// igc_server.hip -- continuous-batching codebook server on one GPU.
//
// Jobs (candidate set + editDist) are submitted from any host thread and
// queued FIFO.  Each GPU pipeline (its own context and stream; up to two, so
// one batch's pair kernel overlaps the previous batch's CSR assembly and
// copies) takes, as soon as it is free, every queued job that fits the HBM
// budget -- the budget is shared by the batches in flight -- and builds their
// conflict graphs in one batched build (igc_batch.hip).  The per-job CSRs
// go to a pool of host threads running the greedy independent set, while the
// pipeline already serves the next batch.  A job's latency is its completion
// time minus its submission time.
//
// Admission estimate per job: symbols + packed words + row pointers + the
// per-string arrays of the row filter and the batch (counts 8 + 8, permutation
// 4, bucket keys 4, set base 4, window items ~3, sorted words) + 48 B per edge
// of the density prior's estimate (edge list 8, col 8, the radix sort's two
// key buffers 32 in the 64-bit form; less in the 32-bit form).  A job whose estimate alone exceeds the budget is refused with
// IGC_ERR_ADMISSION; otherwise FIFO order is kept (a job that does not fit
// waits for memory; nothing behind it overtakes it).
//
// The reference has no server: its clustered loop runs one
// FillAdjListGPU process per cluster on worker threads
// (src/SparseMat.cpp:977-1020, :784-851).
#include "igc_device.hpp"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>
#include <vector>

namespace igc {

struct SJob {
    uint64_t id = 0;
    std::vector<uint8_t> syms;
    uint32_t N = 0, n = 0, ed = 0;
    uint64_t est = 0;
    double t_submit = 0, t_admit = 0, t_gpu = 0, t_done = 0;
    igc_csr csr{};
    uint64_t n_edges = 0;
    std::vector<uint32_t> order;
    int status = IGC_OK;
    bool done = false;
    uint32_t batch = 0, batch_jobs = 0;
};

static uint64_t job_bytes(uint32_t N, uint32_t n, uint32_t ed)
{
    const int K = ed == 0 ? -1 : effective_k(n, ed);
    const uint64_t edges = (K < 0 || N < 2) ? 0 : initial_edge_cap(N, n, K);
    return (uint64_t)N * n + 2ull * N * word_bytes(n) + ((uint64_t)N + 1) * 8 + (uint64_t)N * 32 + edges * 48;
}

} // namespace igc

using namespace igc;

struct igc_server {
    std::vector<igc_ctx *> ctx;
    uint64_t budget = 0, in_use = 0;
    uint32_t max_batch = 512;
    std::mutex mu;
    std::condition_variable cv_queue, cv_greedy, cv_done;
    std::deque<SJob *> queue, greedy_q;
    std::unordered_map<uint64_t, SJob *> jobs;
    std::vector<std::thread> gpu_threads, workers;
    bool stop = false;
    uint32_t gpu_alive = 0;
    uint64_t next_id = 1;
    uint32_t batches = 0;
    uint64_t pairs_done = 0, jobs_done = 0, max_in_use = 0, max_batch_seen = 0;
    double gpu_busy_s = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

static void gpu_loop(igc_server *S, igc_ctx *ctx)
{
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_queue.wait(lk, [&] { return S->stop || !S->queue.empty(); });
        if (S->stop && S->queue.empty()) {
            S->gpu_alive--;
            S->cv_greedy.notify_all();
            return;
        }
        // admit FIFO while the batch fits the shared budget
        std::vector<SJob *> batch;
        uint64_t bytes = 0;
        while (!S->queue.empty() && batch.size() < S->max_batch) {
            SJob *j = S->queue.front();
            if (j->est > S->budget) { // can never fit: refuse
                S->queue.pop_front();
                j->status = set_error(IGC_ERR_ADMISSION, "job %llu needs ~%llu bytes, budget %llu",
                                      (unsigned long long)j->id, (unsigned long long)j->est,
                                      (unsigned long long)S->budget);
                j->done = true;
                j->t_done = S->now();
                S->cv_done.notify_all();
                continue;
            }
            if (S->in_use + bytes + j->est > S->budget) break;
            S->queue.pop_front();
            bytes += j->est;
            batch.push_back(j);
        }
        if (batch.empty()) {
            // memory is held by the other pipeline; wait for it to release
            S->cv_queue.wait(lk);
            continue;
        }
        S->in_use += bytes;
        S->max_in_use = std::max(S->max_in_use, S->in_use);
        S->max_batch_seen = std::max<uint64_t>(S->max_batch_seen, batch.size());
        const uint32_t bno = ++S->batches;
        const double ta = S->now();
        for (SJob *j : batch) {
            j->t_admit = ta;
            j->batch = bno;
            j->batch_jobs = (uint32_t)batch.size();
        }
        lk.unlock();

        std::vector<igc_set> sets(batch.size());
        uint64_t pairs = 0;
        for (size_t t = 0; t < batch.size(); ++t) {
            SJob *j = batch[t];
            sets[t] = igc_set{j->syms.data(), j->N, j->n, j->ed, 0u};
            pairs += (uint64_t)j->N * (j->N > 0 ? j->N - 1 : 0) / 2;
        }
        std::vector<igc_csr> outs(batch.size());
        int rc;
        {
            std::lock_guard<std::mutex> clk(ctx->mu);
            rc = hipSetDevice(ctx->device) == hipSuccess ? batch_build_host_locked(ctx, sets.data(),
                                                                                  (uint32_t)sets.size(), outs.data())
                                                         : IGC_ERR_DEVICE;
        }
        const double tg = S->now();

        lk.lock();
        S->in_use -= bytes;
        S->gpu_busy_s += tg - ta;
        S->pairs_done += pairs;
        for (size_t t = 0; t < batch.size(); ++t) {
            SJob *j = batch[t];
            j->t_gpu = tg;
            if (rc) {
                j->status = rc;
                j->done = true;
                j->t_done = tg;
            } else {
                j->csr = outs[t];
                j->n_edges = outs[t].n_edges;
                j->syms.clear();
                j->syms.shrink_to_fit();
                S->greedy_q.push_back(j);
            }
        }
        S->cv_greedy.notify_all();
        S->cv_done.notify_all();
        S->cv_queue.notify_all(); // memory released: a waiting pipeline may admit now
    }
}

static void greedy_loop(igc_server *S)
{
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // exit only once every GPU pipeline has drained into greedy_q
        S->cv_greedy.wait(lk, [&] { return (S->stop && S->gpu_alive == 0) || !S->greedy_q.empty(); });
        if (S->greedy_q.empty()) return;
        SJob *j = S->greedy_q.front();
        S->greedy_q.pop_front();
        lk.unlock();
        j->order.resize(j->csr.n_nodes);
        uint32_t cnt = 0;
        const int rc = igc_greedy(&j->csr, j->order.data(), &cnt);
        j->order.resize(cnt);
        igc_csr_free(&j->csr);
        lk.lock();
        j->status = rc;
        j->done = true;
        j->t_done = S->now();
        S->jobs_done++;
        S->cv_done.notify_all();
    }
}

static void fill_result(const SJob *j, igc_job_result *out)
{
    std::memset(out, 0, sizeof *out);
    out->id = j->id;
    out->status = j->status;
    out->n_nodes = j->N;
    out->codewords = (uint32_t)j->order.size();
    out->n_edges = j->n_edges;
    out->submit_s = j->t_submit;
    out->admit_s = j->t_admit;
    out->gpu_done_s = j->t_gpu;
    out->done_s = j->t_done;
    out->batch = j->batch;
    out->batch_jobs = j->batch_jobs;
    if (!j->order.empty()) {
        out->order = static_cast<uint32_t *>(std::malloc(sizeof(uint32_t) * j->order.size()));
        if (out->order) std::memcpy(out->order, j->order.data(), sizeof(uint32_t) * j->order.size());
    }
}

extern "C" {

IGC_API int igc_server_create(int device, uint64_t hbm_budget_bytes, uint32_t pipelines, uint32_t host_threads,
                              igc_server **out)
{
    if (!out) return set_error(IGC_ERR_INVALID, "null out");
    *out = nullptr;
    if (pipelines < 1) pipelines = 1;
    if (pipelines > 2) pipelines = 2;
    if (host_threads < 1) host_threads = 1;
    igc_server *S = new (std::nothrow) igc_server();
    if (!S) return set_error(IGC_ERR_OOM, "host allocation failed");
    S->budget = hbm_budget_bytes ? hbm_budget_bytes : (200ull << 30);
    for (uint32_t p = 0; p < pipelines; ++p) {
        igc_ctx *c = nullptr;
        const int rc = igc_ctx_create(device, &c);
        if (rc) {
            for (igc_ctx *x : S->ctx) igc_ctx_destroy(x);
            delete S;
            return rc;
        }
        S->ctx.push_back(c);
    }
    S->gpu_alive = (uint32_t)S->ctx.size();
    for (igc_ctx *c : S->ctx) S->gpu_threads.emplace_back(gpu_loop, S, c);
    for (uint32_t t = 0; t < host_threads; ++t) S->workers.emplace_back(greedy_loop, S);
    *out = S;
    return IGC_OK;
}

IGC_API int igc_server_submit(igc_server *S, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                              uint64_t *job_id)
{
    if (!S || (!syms && N) || !job_id) return set_error(IGC_ERR_INVALID, "null argument");
    if (n < 1 || n > IGC_MAX_LEN) return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%d", n, IGC_MAX_LEN);
    SJob *j = new (std::nothrow) SJob();
    if (!j) return set_error(IGC_ERR_OOM, "host allocation failed");
    try {
        j->syms.assign(syms, syms + (size_t)N * n);
    } catch (...) {
        delete j;
        return set_error(IGC_ERR_OOM, "host allocation failed");
    }
    j->N = N;
    j->n = n;
    j->ed = edit_dist;
    j->est = job_bytes(N, n, edit_dist);
    std::lock_guard<std::mutex> lk(S->mu);
    j->id = S->next_id++;
    j->t_submit = S->now();
    S->jobs[j->id] = j;
    S->queue.push_back(j);
    *job_id = j->id;
    S->cv_queue.notify_one();
    return IGC_OK;
}

IGC_API int igc_server_wait(igc_server *S, uint64_t job_id, int64_t timeout_ms, igc_job_result *out)
{
    if (!S || !out) return set_error(IGC_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return set_error(IGC_ERR_INVALID, "unknown job %llu", (unsigned long long)job_id);
    SJob *j = it->second;
    auto ready = [&] { return j->done; };
    if (timeout_ms < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return IGC_ERR_BUSY;
    fill_result(j, out);
    S->jobs.erase(it);
    delete j;
    return IGC_OK;
}

IGC_API int igc_server_poll(igc_server *S, uint64_t job_id, igc_job_result *out)
{
    return igc_server_wait(S, job_id, 0, out);
}

IGC_API int igc_job_free(igc_job_result *r)
{
    if (r) {
        std::free(r->order);
        r->order = nullptr;
    }
    return IGC_OK;
}

IGC_API double igc_server_clock(igc_server *S) { return S ? S->now() : 0.0; }

IGC_API int igc_server_stats(igc_server *S, igc_server_info *out)
{
    if (!S || !out) return set_error(IGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    out->batches = S->batches;
    out->jobs_done = S->jobs_done;
    out->pairs_done = S->pairs_done;
    out->queued = S->queue.size();
    out->in_use_bytes = S->in_use;
    out->max_in_use_bytes = S->max_in_use;
    out->max_batch_jobs = S->max_batch_seen;
    out->gpu_busy_s = S->gpu_busy_s;
    return IGC_OK;
}

IGC_API int igc_server_destroy(igc_server *S)
{
    if (!S) return IGC_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_queue.notify_all();
    S->cv_greedy.notify_all();
    for (auto &t : S->gpu_threads) t.join();
    for (auto &t : S->workers) t.join();
    for (auto &kv : S->jobs) {
        igc_csr_free(&kv.second->csr);
        delete kv.second;
    }
    for (igc_ctx *c : S->ctx) igc_ctx_destroy(c);
    delete S;
    return IGC_OK;
}

} // extern "C"
