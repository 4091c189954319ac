// igc_host.cpp -- host side of libigc.so: the greedy independent set, neighbour
// order policies and the reference's output formats.  Plain C++17, no device code.
#include "igc_internal.hpp"

#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

namespace igc {

// Blocks of 4 MiB and more come from (and return to) a bounded cache of blocks
// whose pages are already faulted in: a fresh block fills at the page-fault rate
// (~8 GB/s with 8 threads), a cached one at memcpy speed (~30 GB/s), and a
// server or a refinement loop allocates and frees the same sizes over and over.
// The cache holds at most min(16 GiB, physical memory / 8); a block is reused
// for a request of at least half its size.
namespace {
struct HostCache {
    std::mutex mu;
    std::multimap<size_t, void *> idle;       // capacity -> cached block
    std::unordered_map<void *, size_t> lent;  // large blocks handed out -> capacity
    size_t cached = 0, cap = 0;
};
HostCache &host_cache()
{
    static HostCache *c = [] {
        auto *h = new HostCache; // never destroyed: frees may come from other threads at exit
        const long pages = sysconf(_SC_PHYS_PAGES), psz = sysconf(_SC_PAGESIZE);
        const size_t phys = pages > 0 && psz > 0 ? (size_t)pages * (size_t)psz : 0;
        h->cap = std::min<size_t>(16ull << 30, phys / 8);
        return h;
    }();
    return *c;
}
} // namespace

void *host_alloc(size_t bytes)
{
    constexpr size_t kHuge = 2u << 20;
    if (bytes < 2 * kHuge) return std::malloc(bytes);
    const size_t sz = (bytes + kHuge - 1) / kHuge * kHuge;
    HostCache &c = host_cache();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.idle.lower_bound(sz);
        if (it != c.idle.end() && it->first <= 2 * sz) {
            void *p = it->second;
            const size_t cap = it->first;
            c.idle.erase(it);
            c.cached -= cap;
            c.lent[p] = cap;
            return p;
        }
    }
    void *p = std::aligned_alloc(kHuge, sz);
    if (!p) { // give the cached blocks back and try once more
        host_cache_trim();
        p = std::aligned_alloc(kHuge, sz);
    }
    if (!p) return nullptr;
    (void)madvise(p, sz, MADV_HUGEPAGE);
    std::lock_guard<std::mutex> lk(c.mu);
    c.lent[p] = sz;
    return p;
}

void host_free(void *p)
{
    if (!p) return;
    HostCache &c = host_cache();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.lent.find(p);
        if (it != c.lent.end()) {
            const size_t cap = it->second;
            c.lent.erase(it);
            if (c.cached + cap <= c.cap) {
                c.idle.emplace(cap, p);
                c.cached += cap;
                return;
            }
        }
    }
    std::free(p);
}

void host_cache_trim()
{
    HostCache &c = host_cache();
    std::multimap<size_t, void *> drop;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        drop.swap(c.idle);
        c.cached = 0;
    }
    for (auto &b : drop) std::free(b.second);
}

void reorder_threads(igc_csr &csr, uint32_t T)
{
    const uint32_t N = csr.n_nodes;
    for (uint32_t x = 0; x < N; ++x) {
        uint32_t *b = csr.col + csr.row_ptr[x];
        uint32_t *e = csr.col + csr.row_ptr[x + 1];
        std::stable_sort(b, e, [x, T](uint32_t y1, uint32_t y2) {
            const uint32_t m1 = std::min(x, y1), m2 = std::min(x, y2);
            const uint32_t r1 = m1 % T, r2 = m2 % T;
            if (r1 != r2) return r1 < r2;
            if (m1 != m2) return m1 < m2;
            return std::max(x, y1) < std::max(x, y2);
        });
    }
}

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
// share one 8-byte record, and a deleted node's degree field holds kDead, so
// the most frequent test (deleted or degree 0) reads that record only; buckets
// live in one flat array (bucket d never holds more nodes than there are nodes
// of initial degree >= d, since degrees only fall), and the records of the
// next neighbours are prefetched.  The
// visiting order is exactly the reference's.
class Greedy {
  public:
    Greedy(uint32_t N, const uint64_t *rp, const uint32_t *col) : N_(N), rp_(rp), col_(col) {}

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
            std::vector<uint64_t> ge(nb_ + 1, 0);
            for (uint32_t v = 0; v < N_; ++v) ge[st_[v].deg]++;
            for (int64_t d = (int64_t)nb_ - 1; d >= 0; --d) ge[d] += ge[d + 1];
            bk_.resize(nb_);
            uint64_t off = 0;
            for (uint32_t d = 0; d < nb_; ++d) {
                bk_[d].off = off;
                bk_[d].len = 0;
                off += ge[d];
            }
            // capacity 2|E| + N, but only bucket tails are ever touched: leave it
            // uninitialised so untouched pages are never faulted in
            store_.reset(new uint32_t[off > 0 ? off : 1]);
        } catch (const std::bad_alloc &) {
            return set_error(IGC_ERR_OOM, "greedy: host allocation failed");
        }
        tracker_ = nb_;
        for (uint32_t v = 0; v < N_; ++v) {
            push(st_[v].deg, v);
            if (st_[v].deg < tracker_) tracker_ = st_[v].deg;
        }
        uint32_t active = N_, picked = 0;
        try {
            ball_.resize((size_t)maxd + 2);
            live_.resize((size_t)maxd + 1);
        } catch (const std::bad_alloc &) {
            return set_error(IGC_ERR_OOM, "greedy: host allocation failed");
        }
        while (active > 0) {
            const Bucket &b = bk_[tracker_];
            const uint32_t v = store_[b.off + b.len - 1];
            order[picked++] = v;
            // the ball: v's active neighbours in list order, then v (compacted
            // without branches: about half the tests fail, unpredictably)
            const uint64_t v0 = rp_[v], v1 = rp_[v + 1];
            uint32_t nb = 0;
            for (uint64_t q = v0; q < v1; ++q) {
                const uint32_t u = col_[q];
                ball_[nb] = u;
                nb += 1u - dead(u);
            }
            ball_[nb++] = v;
            for (uint32_t bi = 0; bi < nb; ++bi) {
                const uint32_t x = ball_[bi];
                const uint64_t x0 = rp_[x], x1 = rp_[x + 1];
                // x's neighbours that move (active, degree > 0), compacted the same
                // way; a move changes only the moved node's degree, so every test
                // is already fixed before the first move.  An active neighbour of
                // x has degree >= 1 while x is active (a node's degree is its
                // number of active neighbours), so "active" alone decides, read
                // from the bitmap of deleted nodes (N/8 bytes, cache-resident)
                // instead of the node records
                uint32_t m = 0;
                for (uint64_t q = x0; q < x1; ++q) {
                    const uint32_t u = col_[q];
                    live_[m] = u;
                    m += 1u - dead(u);
                }
                for (uint32_t e = 0; e < m; ++e) {
                    // node records kAhead moves ahead, their bucket slots half as far
                    if (e + kAhead < m) __builtin_prefetch(&st_[live_[e + kAhead]]);
                    if (e + kAhead / 2 < m) {
                        const Node nw = st_[live_[e + kAhead / 2]];
                        if (nw.deg < nb_) __builtin_prefetch(&store_[bk_[nw.deg].off + nw.pos]);
                    }
                    const uint32_t u = live_[e];
                    const uint32_t du = st_[u].deg;
                    remove(du, u);
                    st_[u].deg = du - 1;
                    push(du - 1, u);
                    if (du - 1 < tracker_) tracker_ = du - 1;
                }
                remove(st_[x].deg, x);
                st_[x].deg = kDead;
                dead_[x >> 6] |= 1ull << (x & 63);
                --active;
            }
        }
        *count = picked;
        return IGC_OK;
    }

  private:
    static constexpr uint64_t kAhead = 8;
    static constexpr uint32_t kDead = 0xFFFFFFFFu; // degree field of a deleted node
    std::vector<uint32_t> ball_, live_;            // the current ball; a node's moving neighbours
    struct Node {
        uint32_t deg;
        uint32_t pos;
    };
    struct Bucket {
        uint64_t off;
        uint32_t len;
    };
    void push(uint32_t d, uint32_t v)
    {
        Bucket &b = bk_[d];
        st_[v].pos = b.len;
        store_[b.off + b.len++] = v;
    }
    void remove(uint32_t d, uint32_t v)
    {
        Bucket &b = bk_[d];
        const uint32_t p = st_[v].pos;
        const uint32_t last = store_[b.off + b.len - 1];
        store_[b.off + p] = last;
        st_[last].pos = p;
        --b.len;
        if (d == tracker_ && b.len == 0)
            while (tracker_ < nb_ && bk_[tracker_].len == 0) ++tracker_;
    }

    uint32_t N_;
    const uint64_t *rp_;
    const uint32_t *col_;
    uint32_t nb_ = 0, tracker_ = 0;
    std::vector<Node> st_;
    std::vector<uint64_t> dead_; // one bit per deleted node: the neighbour scans test only this
    uint32_t dead(uint32_t u) const { return (uint32_t)(dead_[u >> 6] >> (u & 63)) & 1u; }
    std::vector<Bucket> bk_;
    std::unique_ptr<uint32_t[]> store_;
};

} // namespace igc

using namespace igc;

extern "C" {

IGC_API int igc_greedy(const igc_csr *csr, uint32_t *order, uint32_t *count)
{
    if (!csr || !count || (!order && csr->n_nodes)) return set_error(IGC_ERR_INVALID, "null argument");
    if (csr->location != IGC_ON_HOST) return set_error(IGC_ERR_INVALID, "igc_greedy needs a host CSR");
    if (csr->n_nodes && csr->row_ptr[csr->n_nodes] != 2 * csr->n_edges)
        return set_error(IGC_ERR_INVALID, "igc_greedy needs the symmetric CSR (not an upper-triangle one)");
    *count = 0;
    if (csr->n_nodes == 0) return IGC_OK;
    if (!csr->row_ptr || (!csr->col && csr->n_edges)) return set_error(IGC_ERR_INVALID, "CSR has null arrays");
    Greedy g(csr->n_nodes, csr->row_ptr, csr->col);
    return g.run(order, count);
}

IGC_API int igc_csr_reorder(igc_csr *csr, uint32_t flags)
{
    if (!csr) return set_error(IGC_ERR_INVALID, "null argument");
    if (csr->location != IGC_ON_HOST) return set_error(IGC_ERR_INVALID, "igc_csr_reorder needs a host CSR");
    const uint32_t T = (flags >> 16) & 0xFFFFu;
    if (T > 1) reorder_threads(*csr, T);
    return IGC_OK;
}

IGC_API int igc_write_pair_file(const igc_csr *csr, const char *path)
{
    if (!csr || !path) return set_error(IGC_ERR_INVALID, "null argument");
    if (csr->location != IGC_ON_HOST) return set_error(IGC_ERR_INVALID, "igc_write_pair_file needs a host CSR");
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_error(IGC_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
    std::vector<int32_t> buf;
    buf.reserve(1 << 16);
    bool ok = true;
    for (uint32_t u = 0; u < csr->n_nodes && ok; ++u) {
        for (uint64_t q = csr->row_ptr[u]; q < csr->row_ptr[u + 1]; ++q) {
            const uint32_t v = csr->col[q];
            if (v <= u) continue;
            buf.push_back((int32_t)u);
            buf.push_back((int32_t)v);
            if (buf.size() >= (1u << 16)) {
                ok = std::fwrite(buf.data(), sizeof(int32_t), buf.size(), f) == buf.size();
                buf.clear();
            }
        }
    }
    if (ok && !buf.empty()) ok = std::fwrite(buf.data(), sizeof(int32_t), buf.size(), f) == buf.size();
    if (std::fclose(f) != 0) ok = false;
    if (!ok) return set_error(IGC_ERR_IO, "write to %s failed", path);
    return IGC_OK;
}

IGC_API int igc_write_codebook(const char *dir, const uint8_t *syms, uint32_t n, const uint32_t *order, uint32_t count,
                               const igc_codebook_info *info, char *name_out, size_t name_cap)
{
    if (!dir || !info || (count && (!syms || !order))) return set_error(IGC_ERR_INVALID, "null argument");
    if (n < 1 || n > 99) return set_error(IGC_ERR_INVALID, "bad code length %u", n);
    if (count > 9999999u) return set_error(IGC_ERR_INVALID, "codebook too large for the file name format");
    char name[96];
    std::snprintf(name, sizeof name, "CodeSize-%07u_CodeLen-%02d_MinED-%d.txt", count, info->code_len, info->min_ed);
    std::string path = std::string(dir) + "/" + name;
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return set_error(IGC_ERR_IO, "cannot open %s: %s", path.c_str(), std::strerror(errno));
    // header fields and spacing follow PrintParamsToFile (src/Utils.cpp:311-379)
    std::fprintf(f, "--- Global Parameters ---\n");
    std::fprintf(f, "Code Length:\t\t\t%d\n", info->code_len);
    std::fprintf(f, "Min Codebook Edit Distance:\t%d\n\n", info->min_ed);
    std::fprintf(f, "Max Run:\t\t\t%d\n", info->max_run);
    std::fprintf(f, "Min GC Content:\t\t\t%g\n", info->min_gc);
    std::fprintf(f, "Max GC Content:\t\t\t%g\n\n", info->max_gc);
    std::fprintf(f, "%s\n\n", info->method_line ? info->method_line : "Using Generation Method: External candidates");
    std::fprintf(f, "--- Results Summary ---\n");
    std::fprintf(f, "Number of Candidates:\t\t%lld\n", (long long)info->candidate_num);
    const bool clustered = info->cluster_k > 0;
    if (clustered) {
        std::fprintf(f, "Clustering Method:\t\t%s\n", info->cluster_method ? info->cluster_method : "");
        std::fprintf(f, "Number of Clusters (K):\t\t%d\n", info->cluster_k);
        std::fprintf(f, "Required Identical Iterations:\t%d\n", info->convergence_iterations);
        std::fprintf(f, "Iterations to Converge:\t\t%d\n", info->cluster_iterations);
    } else {
        std::fprintf(f, "Number of Ones in Matrix:\t%lld\n", (long long)info->matrix_ones);
    }
    std::fprintf(f, "Number of Code Words:\t\t%u\n\n", count);
    std::fprintf(f, "--- Performance Metrics ---\n");
    std::fprintf(f, "Number of Threads:\t\t%d\n", info->threads);
    std::fprintf(f, "Candidate Generation Time:\t%.2f\tseconds\n", info->candidates_s);
    if (clustered && info->cluster_iterations > 0) {
        const double it = info->cluster_iterations;
        std::fprintf(f, "Total Clustering Time:\t\t%.2f\tseconds\n", info->fill_s);
        std::fprintf(f, "Avg Clustering Time/Iter:\t%.2f\tseconds\n", info->fill_s / it);
        std::fprintf(f, "Total Solving Time:\t\t%.2f\tseconds\n", info->process_s);
        std::fprintf(f, "Avg Solving Time/Iter:\t\t%.2f\tseconds\n", info->process_s / it);
    } else {
        std::fprintf(f, "Fill Adjacency List Time:\t%.2f\tseconds\n", info->fill_s);
        std::fprintf(f, "Process Matrix Time:\t\t%.2f\tseconds\n", info->process_s);
    }
    std::fprintf(f, "Overall Execution Time:\t\t%.2f\tseconds\n", info->overall_s);
    std::fprintf(f, "=========================================== \n");
    std::vector<char> line(n + 1);
    line[n] = '\n';
    bool ok = true;
    for (uint32_t k = 0; k < count && ok; ++k) {
        const uint8_t *s = syms + (size_t)order[k] * n;
        for (uint32_t p = 0; p < n; ++p) line[p] = (char)('0' + s[p]);
        ok = std::fwrite(line.data(), 1, n + 1, f) == n + 1;
    }
    if (std::fclose(f) != 0) ok = false;
    if (!ok) return set_error(IGC_ERR_IO, "write to %s failed", path.c_str());
    if (name_out && name_cap) {
        std::strncpy(name_out, name, name_cap - 1);
        name_out[name_cap - 1] = '\0';
    }
    return IGC_OK;
}

} // extern "C"
