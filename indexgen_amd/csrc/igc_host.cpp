// igc_host.cpp -- host side of libigc.so: host memory for results, neighbour
// order policies and the reference's output formats (the greedy independent
// set is in igc_greedy.cpp).  Plain C++17, no device code.
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

} // namespace igc

using namespace igc;

extern "C" {

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
