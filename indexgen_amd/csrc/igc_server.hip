// igc_server.hip -- continuous-batching codebook server on one GPU.
//
// Jobs (candidate set + editDist) are submitted from any host thread and
// queued FIFO; their symbols are checked at submission.  Each GPU pipeline
// (its own context and stream; up to four, so one batch's pair kernel overlaps
// another's CSR assembly and copies, and small jobs find a free pipeline while
// others copy the multi-gigabyte CSRs of dense jobs to the host) takes, as
// soon as it is free,
// every queued job that fits the HBM budget and builds their conflict graphs
// in one batched build (igc_batch.hip).  The per-job CSRs go to a pool of host
// threads running the greedy independent set, while the pipeline already
// serves the next batch.  A job's latency is its completion time minus its
// submission time.
//
// HBM budget.  Every device buffer of a pipeline context is charged to the
// server's pool (MemPool): a pipeline reserves max(its buffers, the batch's
// estimate) before a batch, a buffer that grows beyond the reservation extends
// it only while the sum of all reservations stays within the budget, and
// otherwise the allocation fails with IGC_ERR_OOM.  So the device memory the
// pipelines hold never exceeds the budget, also when a job has far more edges
// than estimated.  A batch that runs out is re-run job by job; a job that runs
// out alone is re-run once with every pipeline's buffers freed and the whole
// budget reserved, and fails with IGC_ERR_OOM only then.  After a batch an
// idle pipeline keeps at most its share of the budget in buffers.
//
// Admission estimate per job: symbols + packed words + row pointers + the
// per-string arrays of the row filter and the batch (counts 8 + 8, permutation
// 4, bucket keys 4, set base 4, window items ~3, sorted words) + 48 B per edge
// of the density estimate (edge list 8, col 8, the radix sort's two key
// buffers 32 in the 64-bit form; less in the 32-bit form).  A job whose
// estimate alone exceeds the budget is refused with IGC_ERR_ADMISSION;
// otherwise FIFO order is kept (a job that does not fit waits for memory;
// nothing behind it overtakes it).  Host CSRs waiting for the greedy are
// bounded too: no batch is admitted while they hold more than the host budget.
//
// The reference has no server: its clustered loop runs one
// FillAdjListGPU process per cluster on worker threads
// (src/SparseMat.cpp:977-1020, :784-851).
#include "igc_device.hpp"

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <unistd.h>
#include <thread>
#include <unordered_map>
#include <vector>

namespace igc {

struct SJob {
    uint64_t id = 0;
    std::vector<uint8_t> syms;
    uint32_t N = 0, n = 0, ed = 0;
    // a refinement job (igc_server_submit_refine): the clustered loop's
    // parameters; it runs alone on a pipeline, its state on the device
    bool refine = false;
    uint32_t k = 0, km_iters = 0, max_rounds = 0, converge_n = 0;
    uint64_t seed = 0;
    uint64_t est = 0;
    double t_submit = 0, t_admit = 0, t_gpu = 0, t_greedy = 0, t_done = 0;
    igc_csr csr{};
    uint64_t n_edges = 0, host_bytes = 0;
    std::vector<uint32_t> order;
    int status = IGC_OK;
    std::string err; // message of a failed job, handed to the waiting thread
    bool done = false;
    bool solo = false;    // admitted alone (after its batch ran out of budget)
    bool retried = false; // already re-run with the whole budget
    uint32_t batch = 0, batch_jobs = 0;
};

static uint64_t job_bytes(uint32_t N, uint32_t n, uint32_t ed)
{
    const int K = ed == 0 ? -1 : effective_k(n, ed);
    const uint64_t edges = (K < 0 || N < 2) ? 0 : initial_edge_cap(N, n, K);
    return (uint64_t)N * n + 2ull * N * word_bytes(n) + ((uint64_t)N + 1) * 8 + (uint64_t)N * 32 + edges * 48;
}

// A refinement job's reservation: the resident strings (set, round, cluster
// order: 3 B per symbol), index lists, the K-means' packed words and
// assignment, and the first round's batched build (its clusters' pairs,
// about N^2 / 2k, at the density estimate).
static uint64_t refine_bytes(uint32_t N, uint32_t n, uint32_t ed, uint32_t k)
{
    const uint32_t kk = std::max<uint32_t>(1, std::min<uint32_t>(k, std::max<uint32_t>(N, 1)));
    const uint32_t per = std::max<uint32_t>(2, N / kk);
    return 3ull * N * n + 24ull * N + (uint64_t)kk * job_bytes(per, n, ed);
}

// Host memory the CSRs waiting for the greedy may hold before the pipelines
// stop admitting: a quarter of physical memory, at most 64 GiB.
static uint64_t default_host_budget()
{
    const long pages = sysconf(_SC_PHYS_PAGES), psz = sysconf(_SC_PAGE_SIZE);
    const uint64_t phys = pages > 0 && psz > 0 ? (uint64_t)pages * (uint64_t)psz : (64ull << 30);
    return std::max<uint64_t>(1ull << 30, std::min<uint64_t>(64ull << 30, phys / 4));
}

} // namespace igc

using namespace igc;

struct igc_server {
    std::vector<igc_ctx *> ctx;
    std::vector<uint8_t> busy; // per pipeline: a batch is running
    MemPool pool;              // the HBM budget, charged by every pipeline context's buffers
    uint64_t budget = 0;
    uint64_t host_budget = 0, host_pending = 0; // host CSR bytes waiting for the greedy
    int exclusive = -1;        // pipeline re-running a job with the whole budget
    uint32_t max_batch = 512;
    uint32_t refine_threads = 8; // host greedy threads of a refinement job
    std::mutex mu;
    std::condition_variable cv_queue, cv_greedy, cv_done;
    std::deque<SJob *> queue, greedy_q;
    std::unordered_map<uint64_t, SJob *> jobs;
    std::vector<std::thread> gpu_threads, workers;
    bool stop = false;
    uint32_t gpu_alive = 0;
    uint64_t next_id = 1;
    uint32_t batches = 0;
    uint64_t pairs_done = 0, jobs_done = 0, max_batch_seen = 0;
    double gpu_busy_s = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

static void finish_job(igc_server *S, SJob *j, int status, const char *msg)
{
    j->status = status;
    if (status) j->err = msg ? msg : "";
    j->done = true;
    j->t_done = S->now();
}

// Frees the buffers of every idle pipeline other than p (S->mu held; an idle
// pipeline holds no context lock).  Returns whether any memory was released.
static bool trim_idle_others(igc_server *S, uint32_t p)
{
    bool freed = false;
    for (uint32_t q = 0; q < S->ctx.size(); ++q) {
        if (q == p || S->busy[q] || S->ctx[q]->mem.used == 0) continue;
        std::lock_guard<std::mutex> clk(S->ctx[q]->mu);
        ctx_trim_locked(S->ctx[q]);
        freed = true;
    }
    return freed;
}

// A refinement job on pipeline p's context (S->mu not held): the whole loop,
// its strings and round state staying on the device; the per-cluster greedy
// runs on `threads` host threads.
static int run_refine(igc_ctx *ctx, SJob *j, uint32_t threads, std::string &msg)
{
    std::lock_guard<std::mutex> clk(ctx->mu);
    int rc = hipSetDevice(ctx->device) == hipSuccess ? IGC_OK : set_error(IGC_ERR_DEVICE, "hipSetDevice failed");
    if (!rc) {
        j->order.resize(std::max<uint32_t>(j->N, 1));
        uint32_t cnt = 0;
        rc = refine_locked(ctx, j->syms.data(), j->N, j->n, j->ed, j->k, j->km_iters, j->max_rounds, j->converge_n,
                           j->seed, threads, ctx->graph_replay, j->order.data(), &cnt, nullptr);
        j->order.resize(rc ? 0 : cnt);
    }
    if (rc) msg = last_error();
    return rc;
}

// One batched build of the jobs on pipeline p's context (S->mu not held).
static int run_batch(igc_server *S, igc_ctx *ctx, const std::vector<SJob *> &batch, std::vector<igc_csr> &outs,
                     std::string &msg)
{
    std::vector<igc_set> sets(batch.size());
    for (size_t t = 0; t < batch.size(); ++t) sets[t] = igc_set{batch[t]->syms.data(), batch[t]->N, batch[t]->n,
                                                                batch[t]->ed, 0u};
    outs.assign(batch.size(), igc_csr{});
    std::lock_guard<std::mutex> clk(ctx->mu);
    int rc = hipSetDevice(ctx->device) == hipSuccess
                 ? batch_build_host_locked(ctx, sets.data(), (uint32_t)sets.size(), outs.data())
                 : set_error(IGC_ERR_DEVICE, "hipSetDevice(%d) failed", ctx->device);
    if (rc) msg = last_error();
    return rc;
}

static void gpu_loop(igc_server *S, uint32_t p)
{
    igc_ctx *ctx = S->ctx[p];
    MemSlot &slot = ctx->mem;
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // admit when jobs wait, no pipeline runs alone, and the host CSRs
        // waiting for the greedy leave room (a destroy drains the queue first)
        S->cv_queue.wait(lk, [&] {
            return (S->stop && S->queue.empty()) ||
                   (!S->queue.empty() && S->exclusive < 0 &&
                    (S->host_pending <= S->host_budget || S->host_pending == 0));
        });
        if (S->queue.empty()) {
            S->gpu_alive--;
            S->cv_greedy.notify_all();
            return;
        }
        // admit FIFO while the batch's estimate fits next to what the other
        // pipelines have reserved
        std::vector<SJob *> batch;
        uint64_t bytes = 0;
        while (!S->queue.empty() && batch.size() < S->max_batch) {
            SJob *j = S->queue.front();
            if (j->est > S->budget) { // can never fit: refuse
                S->queue.pop_front();
                char m[160];
                std::snprintf(m, sizeof m, "job %llu needs ~%llu bytes of HBM, the budget is %llu",
                              (unsigned long long)j->id, (unsigned long long)j->est, (unsigned long long)S->budget);
                finish_job(S, j, IGC_ERR_ADMISSION, m);
                S->cv_done.notify_all();
                continue;
            }
            if ((j->solo || j->refine) && !batch.empty()) break;
            if (!slot.reserve(bytes + j->est)) break;
            S->queue.pop_front();
            bytes += j->est;
            batch.push_back(j);
            if (j->solo || j->refine) break;
        }
        if (batch.empty()) {
            if (S->queue.empty()) continue;
            // the head job does not fit next to the other pipelines' buffers:
            // free idle ones, else wait for a running batch to end
            if (trim_idle_others(S, p)) continue;
            bool busy_other = false;
            for (uint32_t q = 0; q < S->ctx.size(); ++q) busy_other |= q != p && S->busy[q];
            if (!busy_other && slot.used) {
                std::lock_guard<std::mutex> clk(ctx->mu);
                ctx_trim_locked(ctx);
                continue;
            }
            S->cv_queue.wait_for(lk, std::chrono::milliseconds(busy_other ? 1000 : 10));
            continue;
        }
        S->busy[p] = 1;
        S->max_batch_seen = std::max<uint64_t>(S->max_batch_seen, batch.size());
        const uint32_t bno = ++S->batches;
        const double ta = S->now();
        uint64_t pairs = 0;
        for (SJob *j : batch) {
            j->t_admit = ta;
            j->batch = bno;
            j->batch_jobs = (uint32_t)batch.size();
            pairs += (uint64_t)j->N * (j->N > 0 ? j->N - 1 : 0) / 2;
        }
        lk.unlock();
        std::vector<igc_csr> outs;
        std::string msg;
        if (batch[0]->refine) {
            // the whole clustered loop on this pipeline; the result is final
            SJob *j = batch[0];
            const int rc = run_refine(ctx, j, S->refine_threads, msg);
            lk.lock();
            S->busy[p] = 0;
            const double tg = S->now();
            S->gpu_busy_s += tg - ta;
            j->t_gpu = j->t_greedy = tg;
            j->syms.clear();
            j->syms.shrink_to_fit();
            finish_job(S, j, rc, msg.c_str());
            if (!rc) S->jobs_done++;
            if (slot.used > S->budget / S->ctx.size()) {
                std::lock_guard<std::mutex> clk(ctx->mu);
                ctx_trim_locked(ctx);
            }
            slot.settle();
            S->cv_done.notify_all();
            S->cv_queue.notify_all();
            continue;
        }
        int rc = run_batch(S, ctx, batch, outs, msg);
        lk.lock();
        S->busy[p] = 0;
        if (rc == IGC_ERR_OOM && batch.size() > 1) {
            // over the budget: run these jobs one by one, in their order, ahead of the queue
            for (auto it = batch.rbegin(); it != batch.rend(); ++it) {
                (*it)->solo = true;
                S->queue.push_front(*it);
            }
        } else if (rc == IGC_ERR_OOM && !batch[0]->retried) {
            // one job over the budget next to the other pipelines: run it again
            // with every pipeline's buffers freed and the whole budget reserved
            SJob *j = batch[0];
            j->retried = true;
            S->cv_queue.wait(lk, [&] { return S->exclusive < 0; });
            S->exclusive = (int)p;
            S->cv_queue.wait(lk, [&] {
                for (uint32_t q = 0; q < S->ctx.size(); ++q)
                    if (q != p && S->busy[q]) return false;
                return true;
            });
            trim_idle_others(S, p);
            {
                std::lock_guard<std::mutex> clk(ctx->mu);
                ctx_trim_locked(ctx);
            }
            (void)slot.reserve(S->budget);
            S->busy[p] = 1;
            lk.unlock();
            rc = run_batch(S, ctx, batch, outs, msg);
            lk.lock();
            S->busy[p] = 0;
            S->exclusive = -1;
        }
        const bool requeued = rc == IGC_ERR_OOM && batch.size() > 1;
        const double tg = S->now();
        if (!requeued) {
            S->gpu_busy_s += tg - ta;
            if (!rc) S->pairs_done += pairs;
            for (size_t t = 0; t < batch.size(); ++t) {
                SJob *j = batch[t];
                j->t_gpu = tg;
                if (rc) {
                    finish_job(S, j, rc, msg.c_str());
                } else {
                    j->csr = outs[t];
                    j->n_edges = outs[t].n_edges;
                    j->host_bytes = 8ull * ((uint64_t)j->N + 1) + 8ull * outs[t].n_edges;
                    S->host_pending += j->host_bytes;
                    j->syms.clear();
                    j->syms.shrink_to_fit();
                    S->greedy_q.push_back(j);
                }
            }
        }
        // an idle pipeline keeps at most its share of the budget in buffers
        if (slot.used > S->budget / S->ctx.size()) {
            std::lock_guard<std::mutex> clk(ctx->mu);
            ctx_trim_locked(ctx);
        }
        slot.settle();
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
        j->t_greedy = S->now();
        lk.unlock();
        j->order.resize(j->csr.n_nodes);
        uint32_t cnt = 0;
        const int rc = igc_greedy(&j->csr, j->order.data(), &cnt);
        const std::string msg = rc ? std::string(last_error()) : std::string();
        j->order.resize(cnt);
        igc_csr_free(&j->csr);
        lk.lock();
        S->host_pending -= std::min(S->host_pending, j->host_bytes);
        finish_job(S, j, rc, msg.c_str());
        S->jobs_done++;
        S->cv_done.notify_all();
        S->cv_queue.notify_all(); // host memory released
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
    out->greedy_start_s = j->t_greedy;
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
    if (pipelines > 4) pipelines = 4;
    if (host_threads < 1) host_threads = 1;
    igc_server *S = new (std::nothrow) igc_server();
    if (!S) return set_error(IGC_ERR_OOM, "host allocation failed");
    S->budget = hbm_budget_bytes ? hbm_budget_bytes : (200ull << 30);
    S->pool.limit = S->budget;
    S->host_budget = default_host_budget();
    for (uint32_t p = 0; p < pipelines; ++p) {
        igc_ctx *c = nullptr;
        const int rc = igc_ctx_create(device, &c);
        if (rc) {
            for (igc_ctx *x : S->ctx) igc_ctx_destroy(x);
            delete S;
            return rc;
        }
        ctx_attach_pool(c, &S->pool);
        S->ctx.push_back(c);
    }
    S->busy.assign(S->ctx.size(), 0);
    S->gpu_alive = (uint32_t)S->ctx.size();
    for (uint32_t p = 0; p < S->ctx.size(); ++p) S->gpu_threads.emplace_back(gpu_loop, S, p);
    S->refine_threads = host_threads;
    for (uint32_t t = 0; t < host_threads; ++t) S->workers.emplace_back(greedy_loop, S);
    *out = S;
    return IGC_OK;
}

IGC_API int igc_server_submit(igc_server *S, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                              uint64_t *job_id)
{
    if (!S || (!syms && N) || !job_id) return set_error(IGC_ERR_INVALID, "null argument");
    if (n < 1 || n > IGC_MAX_LEN) return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%d", n, IGC_MAX_LEN);
    // symbols are checked here, so one bad job cannot fail the batch it would share
    const size_t bytes = (size_t)N * n;
    const uint8_t top = n > IGC_MAX_QUATERNARY_LEN ? 1 : 3;
    uint8_t any = 0;
    for (size_t b = 0; b < bytes; ++b) any |= (uint8_t)(syms[b] > top);
    if (any)
        return set_error(IGC_ERR_INVALID, n > IGC_MAX_QUATERNARY_LEN
                                              ? "symbol code above 1 in a binary string (strings longer than %d "
                                                "take symbols 0/1 only)"
                                              : "symbol code above 3 in input (strings up to %d)",
                         IGC_MAX_QUATERNARY_LEN);
    SJob *j = new (std::nothrow) SJob();
    if (!j) return set_error(IGC_ERR_OOM, "host allocation failed");
    try {
        j->syms.assign(syms, syms + bytes);
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

IGC_API int igc_server_submit_refine(igc_server *S, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                                     uint32_t k, uint32_t kmeans_iters, uint32_t max_rounds, uint32_t converge_n,
                                     uint64_t seed, uint64_t *job_id)
{
    if (!S || (!syms && N) || !job_id) return set_error(IGC_ERR_INVALID, "null argument");
    if (n < 1 || n > IGC_MAX_QUATERNARY_LEN)
        return set_error(IGC_ERR_INVALID, "string length n=%u outside 1..%d (K-means runs on quaternary strings)", n,
                         IGC_MAX_QUATERNARY_LEN);
    if (k < 1) return set_error(IGC_ERR_INVALID, "k must be >= 1");
    const size_t bytes = (size_t)N * n;
    uint8_t any = 0;
    for (size_t b = 0; b < bytes; ++b) any |= (uint8_t)(syms[b] > 3);
    if (any) return set_error(IGC_ERR_INVALID, "symbol code above 3 in input");
    SJob *j = new (std::nothrow) SJob();
    if (!j) return set_error(IGC_ERR_OOM, "host allocation failed");
    try {
        j->syms.assign(syms, syms + bytes);
    } catch (...) {
        delete j;
        return set_error(IGC_ERR_OOM, "host allocation failed");
    }
    j->N = N;
    j->n = n;
    j->ed = edit_dist;
    j->refine = true;
    j->k = k;
    j->km_iters = kmeans_iters;
    j->max_rounds = max_rounds;
    j->converge_n = converge_n;
    j->seed = seed;
    j->est = refine_bytes(N, n, edit_dist, k);
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
    if (j->status) set_error(j->status, "%s", j->err.c_str()); // the job's error, for igc_last_error()
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
    {
        std::lock_guard<std::mutex> plk(S->pool.mu);
        out->in_use_bytes = S->pool.charged;
        out->max_in_use_bytes = S->pool.peak;
    }
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
