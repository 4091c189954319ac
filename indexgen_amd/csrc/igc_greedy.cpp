// igc_greedy.cpp -- min-degree greedy independent set of a host CSR, with the
// bucket mechanics of the reference's AdjList (src/SparseMat.cpp:81-251): one
// thread makes every bucket move in the reference's order; for large graphs,
// helper threads compact the neighbour lists it will walk ahead of it.  Plain
// C++17, no device code.
#include "igc_internal.hpp"

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

namespace igc {

// Min-degree greedy over a CSR, with the bucket mechanics of the reference's
// AdjList (src/SparseMat.cpp:81-251):
//   - bucket d lists the active nodes of current degree d; built in node order
//     (RowsBySum, :92-120);
//   - the pick is the LAST node of the lowest non-empty bucket (MinSumRow, :122-127);
//   - DelBall (:229-251) deletes the pick's active neighbours in list order and
//     then the pick; each deletion (DelRowCol, :207-227) moves every active
//     neighbour u from bucket deg(u) to deg(u)-1 (DecreaseSum, :156-167): a
//     swap-remove that fills the hole with the bucket's last node (DeleteRow,
//     :139-154) and an append to the lower bucket.
// Layout for speed, not for semantics: a node's degree and bucket position
// share one 8-byte record, and a deleted node's degree field holds kDead;
// each bucket is its own compact array; the neighbour scans test a bitmap of
// deleted nodes.
//
// A pick is one sequence of bucket operations: for each ball member x in
// order, a move of each of x's neighbours still active at that point (in x's
// list order), then the removal of x.  Which neighbours are active at x is
// known once the ball is (the nodes deleted before the pick and the ball
// members before x are not), so with helper threads the scans that find them
// run ahead of the moves (scanning ahead, below), and the moves themselves
// stay on one thread, in order.
class Greedy {
  public:
    Greedy(uint32_t N, const uint64_t *rp, const uint32_t *col, uint32_t threads)
        : N_(N), rp_(rp), col_(col), T_(std::max<uint32_t>(threads, 1))
    {
    }

    int run(uint32_t *order, uint32_t *count)
    {
        uint32_t maxd = 0;
        try {
            st_.resize(N_);
            dead_.assign(((size_t)N_ + 63) / 64, 0ull);
            for (uint32_t v = 0; v < N_; ++v) {
                st_[v].deg = (uint32_t)(rp_[v + 1] - rp_[v]);
                maxd = std::max(maxd, st_[v].deg);
            }
            nb_ = maxd + 1;
            // each bucket starts with room for its nodes of that initial degree
            std::vector<uint32_t> cnt(nb_, 0);
            for (uint32_t v = 0; v < N_; ++v) cnt[st_[v].deg]++;
            bk_.assign(nb_, Bucket{});
            for (uint32_t d = 0; d < nb_; ++d)
                if (cnt[d]) reserve(bk_[d], cnt[d]);
            ball_.resize((size_t)maxd + 2);
            live_.resize((size_t)maxd + 1);
            if (T_ > 1) ahead_setup();
        } catch (const std::bad_alloc &) {
            return set_error(IGC_ERR_OOM, "greedy: host allocation failed");
        }
        tracker_ = nb_;
        uint32_t active = N_, picked = 0;
        try { // a bucket that grows past its room can fail to allocate
        for (uint32_t v = 0; v < N_; ++v) {
            push(st_[v].deg, v);
            if (st_[v].deg < tracker_) tracker_ = st_[v].deg;
        }
        while (active > 0) {
            const Bucket &b = bk_[tracker_];
            const uint32_t v = b.p[b.len - 1];
            order[picked++] = v;
            // the ball: v's active neighbours in list order, then v (compacted
            // without branches: about half the tests fail, unpredictably)
            const uint64_t v0 = rp_[v], v1 = rp_[v + 1];
            uint32_t nb = 0;
            uint64_t work = v1 - v0;
            for (uint64_t q = v0; q < v1; ++q) {
                const uint32_t u = col_[q];
                ball_[nb] = u;
                nb += 1u - dead(u);
            }
            ball_[nb++] = v;
            if (!helpers_.empty()) {
                for (uint32_t bi = 0; bi + 1 < nb; ++bi) work += rp_[ball_[bi] + 1] - rp_[ball_[bi]];
            }
            if (!helpers_.empty() && work >= kAheadMin) pick_ahead(nb);
            else pick_one(nb);
            active -= nb;
        }
        } catch (const std::bad_alloc &) {
            return set_error(IGC_ERR_OOM, "greedy: host allocation failed");
        }
        if (!helpers_.empty()) ahead_stop();
        *count = picked;
        return IGC_OK;
    }
    ~Greedy()
    {
        if (!helpers_.empty()) ahead_stop();
        for (Bucket &b : bk_) std::free(b.p);
    }

  private:
    static constexpr uint64_t kAhead = 8;
    static constexpr uint32_t kDead = 0xFFFFFFFFu; // degree field of a deleted node
    static constexpr uint64_t kAheadMin = 1u << 15; // neighbours scanned by a pick worth the helpers
    struct Node {
        uint32_t deg;
        uint32_t pos;
    };
    // A bucket's nodes in their own array, grown by doubling: the buckets of a
    // pick stay within a few hundred pages (a flat array with every bucket at
    // its largest possible size put each bucket on pages of its own, and the
    // moves of a 400,000-node graph missed the TLB on most of them).
    struct Bucket {
        uint32_t *p = nullptr;
        uint32_t len = 0, cap = 0;
    };
    static void reserve(Bucket &b, uint32_t cap)
    {
        uint32_t *q = static_cast<uint32_t *>(std::realloc(b.p, sizeof(uint32_t) * (size_t)cap));
        if (!q) throw std::bad_alloc();
        b.p = q;
        b.cap = cap;
    }
    void push(uint32_t d, uint32_t v)
    {
        Bucket &b = bk_[d];
        if (b.len == b.cap) reserve(b, b.cap ? 2 * b.cap : 16);
        st_[v].pos = b.len;
        b.p[b.len++] = v;
    }
    void remove(uint32_t d, uint32_t v)
    {
        Bucket &b = bk_[d];
        const uint32_t p = st_[v].pos;
        const uint32_t last = b.p[b.len - 1];
        b.p[p] = last;
        st_[last].pos = p;
        --b.len;
        if (d == tracker_ && b.len == 0)
            while (tracker_ < nb_ && bk_[tracker_].len == 0) ++tracker_;
    }
    uint32_t dead(uint32_t u) const { return (uint32_t)(dead_[u >> 6] >> (u & 63)) & 1u; }

    // The pick's bucket operations on this thread, in order.
    void pick_one(uint32_t nb)
    {
        for (uint32_t bi = 0; bi < nb; ++bi) {
            const uint32_t x = ball_[bi];
            // x's neighbours that move (active, degree > 0), compacted the same
            // way; a move changes only the moved node's degree, so every test
            // is already fixed before the first move.  An active neighbour of
            // x has degree >= 1 while x is active (a node's degree is its
            // number of active neighbours), so "active" alone decides
            uint32_t m = 0;
            for (uint64_t q = rp_[x], q1 = rp_[x + 1]; q < q1; ++q) {
                const uint32_t u = col_[q];
                live_[m] = u;
                m += 1u - dead(u);
            }
            move_live(live_.data(), m);
            remove(st_[x].deg, x);
            st_[x].deg = kDead;
            dead_[x >> 6] |= 1ull << (x & 63);
        }
    }

    // Moves each listed node one bucket down, in order.
    void move_live(const uint32_t *live, uint32_t m)
    {
        for (uint32_t e = 0; e < m; ++e) {
            // node records kAhead moves ahead, their bucket slots half as far
            if (e + kAhead < m) __builtin_prefetch(&st_[live[e + kAhead]]);
            if (e + kAhead / 2 < m) {
                const Node nw = st_[live[e + kAhead / 2]];
                if (nw.deg < nb_) __builtin_prefetch(&bk_[nw.deg].p[nw.pos]);
            }
            const uint32_t u = live[e];
            const uint32_t du = st_[u].deg;
            remove(du, u);
            st_[u].deg = du - 1;
            push(du - 1, u);
            if (du - 1 < tracker_) tracker_ = du - 1;
        }
    }

    // ---- scanning ahead ----
    // Helper h (0..H-1) compacts the live neighbours of ball members h, h+H,
    // h+2H, ... into the member's slice of buf_ (at the prefix of the ball's
    // row lengths) and publishes each list's length with the pick's stamp.  A
    // member's live neighbours are those not deleted before the pick and not
    // an earlier ball member; each helper keeps its own bitmap of those, adding
    // the members it passes.  The calling thread makes the moves of member bi
    // as soon as its list is published, in ball order, so the selection is the
    // one-thread selection.
    void ahead_setup()
    {
        const uint32_t H = T_ - 1;
        hdead_.assign(H, std::vector<uint64_t>(dead_.size(), 0ull));
        dead0_.assign(dead_.size(), 0ull);
        ready_.reset(new std::atomic<uint64_t>[ball_.size()]);
        for (size_t k = 0; k < ball_.size(); ++k) ready_[k].store(0, std::memory_order_relaxed);
        wsum_.resize(ball_.size() + 1);
        for (uint32_t h = 0; h < H; ++h) helpers_.emplace_back([this, h] { helper(h); });
    }
    void ahead_stop()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        for (auto &t : helpers_) t.join();
        helpers_.clear();
    }
    void helper(uint32_t h)
    {
        const uint32_t H = T_ - 1;
        uint64_t seen = 0;
        for (;;) {
            uint64_t g = gen_.load(std::memory_order_acquire);
            if (g == seen) {
                const auto t0 = std::chrono::steady_clock::now();
                for (uint32_t spin = 1; g == seen; ++spin) {
                    _mm_pause();
                    g = gen_.load(std::memory_order_acquire);
                    if ((spin & 255) == 0 && std::chrono::steady_clock::now() - t0 > kSpin) break;
                }
            }
            if (g == seen) {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_.load(std::memory_order_acquire) != seen; });
                g = gen_.load(std::memory_order_acquire);
            }
            seen = g;
            if (stop_) return;
            const uint32_t nb = ahead_nb_, stamp = ahead_stamp_;
            uint64_t *cd = hdead_[h].data();
            std::copy(dead0_.begin(), dead0_.end(), cd);
            uint32_t marked = 0; // ball members 0..marked-1 are in cd
            for (uint32_t bi = h; bi < nb; bi += H) {
                for (; marked < bi; ++marked) {
                    const uint32_t x = ball_[marked];
                    cd[x >> 6] |= 1ull << (x & 63);
                }
                const uint32_t x = ball_[bi];
                uint32_t *live = buf_.data() + wsum_[bi];
                uint32_t m = 0;
                for (uint64_t q = rp_[x], q1 = rp_[x + 1]; q < q1; ++q) {
                    const uint32_t u = col_[q];
                    live[m] = u;
                    m += 1u - (uint32_t)((cd[u >> 6] >> (u & 63)) & 1u);
                }
                ready_[bi].store(((uint64_t)stamp << 32) | m, std::memory_order_release);
            }
        }
    }

    // One pick with the scans done by the helpers.
    void pick_ahead(uint32_t nb)
    {
        wsum_[0] = 0;
        for (uint32_t bi = 0; bi < nb; ++bi) wsum_[bi + 1] = wsum_[bi] + (rp_[ball_[bi] + 1] - rp_[ball_[bi]]);
        if (buf_.size() < wsum_[nb]) buf_.resize(wsum_[nb] + wsum_[nb] / 4);
        dead0_ = dead_; // the helpers' start (this thread updates dead_ meanwhile)
        ahead_nb_ = nb;
        ++ahead_stamp_;
        {
            std::lock_guard<std::mutex> lk(mu_);
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        for (uint32_t bi = 0; bi < nb; ++bi) {
            uint64_t r = ready_[bi].load(std::memory_order_acquire);
            for (uint32_t spin = 0; (uint32_t)(r >> 32) != ahead_stamp_; ++spin) {
                _mm_pause();
                if (spin > (1u << 16)) std::this_thread::yield();
                r = ready_[bi].load(std::memory_order_acquire);
            }
            const uint32_t x = ball_[bi];
            move_live(buf_.data() + wsum_[bi], (uint32_t)r);
            remove(st_[x].deg, x);
            st_[x].deg = kDead;
            dead_[x >> 6] |= 1ull << (x & 63);
        }
    }

    uint32_t N_;
    const uint64_t *rp_;
    const uint32_t *col_;
    uint32_t T_;
    uint32_t nb_ = 0, tracker_ = 0;
    std::vector<Node> st_;
    std::vector<uint64_t> dead_; // one bit per deleted node: the neighbour scans test only this
    std::vector<Bucket> bk_;
    std::vector<uint32_t> ball_, live_; // the current ball; a node's moving neighbours
    // scanning ahead
    static constexpr std::chrono::microseconds kSpin{300};
    std::vector<std::thread> helpers_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::atomic<uint64_t> gen_{0};
    bool stop_ = false;
    uint32_t ahead_nb_ = 0, ahead_stamp_ = 0;
    std::unique_ptr<std::atomic<uint64_t>[]> ready_; // per ball member: stamp << 32 | live count
    std::vector<std::vector<uint64_t>> hdead_;      // per helper: deleted-node bitmap
    std::vector<uint64_t> dead0_;                   // deleted nodes before the pick
    std::vector<uint32_t> buf_;                     // live lists of the ball members
    std::vector<uint64_t> wsum_;                    // their offsets
};

// Threads for a graph: one below 2^24 edges (a few tens of milliseconds on one
// thread), else the moving thread and two helpers, which keep ahead of the
// moves (more helpers measured no faster).
static uint32_t auto_threads(uint64_t edges)
{
    if (edges < (1ull << 24)) return 1;
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    return std::min(hw, 3u);
}

} // namespace igc

using namespace igc;

extern "C" {

IGC_API int igc_greedy_mt(const igc_csr *csr, uint32_t *order, uint32_t *count, uint32_t threads)
{
    if (!csr || !count || (!order && csr->n_nodes)) return set_error(IGC_ERR_INVALID, "null argument");
    if (csr->location != IGC_ON_HOST) return set_error(IGC_ERR_INVALID, "igc_greedy needs a host CSR");
    if (csr->n_nodes && csr->row_ptr[csr->n_nodes] != 2 * csr->n_edges)
        return set_error(IGC_ERR_INVALID, "igc_greedy needs the symmetric CSR (not an upper-triangle one)");
    *count = 0;
    if (csr->n_nodes == 0) return IGC_OK;
    if (!csr->row_ptr || (!csr->col && csr->n_edges)) return set_error(IGC_ERR_INVALID, "CSR has null arrays");
    const uint32_t T = threads ? std::min<uint32_t>(threads, 17) : auto_threads(csr->n_edges);
    Greedy g(csr->n_nodes, csr->row_ptr, csr->col, T);
    return g.run(order, count);
}

IGC_API int igc_greedy(const igc_csr *csr, uint32_t *order, uint32_t *count)
{
    return igc_greedy_mt(csr, order, count, 0);
}

} // extern "C"
