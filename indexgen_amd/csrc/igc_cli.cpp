// igc_cli.cpp -- a C++17 host on the C ABI, following the reference's
// CodebookAdjList flow (src/SparseMat.cpp:734-777): fill the adjacency on the
// GPU, run the greedy codebook, write the codebook file (ToFile,
// src/Utils.cpp:394-418), optionally verify (VerifyDist GPU mode, :203-288).
//
//   igc_cli --cands FILE --min-ed D [--order-threads T] [--out DIR] [--verify]
//           [--pair-file PATH] [--refine K --kmeans-iters I --rounds R --converge C
//            --cluster-method NAME]
//   igc_cli --cands FILE --min-ed D --pair-file PATH --graph-only
//
// --graph-only does what the reference's per-call GPU script does
// (src/gpu_graph_generator.py, run by FillAdjListGPU, src/SparseMat.cpp:482-533):
// candidates file in, LE int32 pair file out, no greedy and no codebook.
//
// FILE holds one candidate per line ('0'..'3' or A,C,G,T), all of one length.
#include "igc.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

static int die(const char *what)
{
    std::cerr << what << ": " << igc_last_error() << std::endl;
    return 1;
}

static int sym_code(char c)
{
    switch (c) {
    case '0': case 'A': return 0;
    case '1': case 'C': return 1;
    case '2': case 'G': return 2;
    case '3': case 'T': return 3;
    default: return -1;
    }
}

int main(int argc, char **argv)
{
    std::string cands, out = ".", pair_file;
    int min_ed = -1, order_t = 1;
    bool verify = false, graph_only = false;
    uint32_t refine_k = 0, km_iters = 10, rounds = 5, converge = 3;
    std::string cluster_method = "kmeans_fp32";
    for (int a = 1; a < argc; ++a) {
        std::string s = argv[a];
        auto next = [&]() -> const char * { return a + 1 < argc ? argv[++a] : ""; };
        if (s == "--cands") cands = next();
        else if (s == "--min-ed") min_ed = std::atoi(next());
        else if (s == "--order-threads") order_t = std::atoi(next());
        else if (s == "--out") out = next();
        else if (s == "--verify") verify = true;
        else if (s == "--graph-only") graph_only = true;
        else if (s == "--pair-file") pair_file = next();
        else if (s == "--refine") refine_k = (uint32_t)std::atoi(next());
        else if (s == "--kmeans-iters") km_iters = (uint32_t)std::atoi(next());
        else if (s == "--rounds") rounds = (uint32_t)std::atoi(next());
        else if (s == "--converge") converge = (uint32_t)std::atoi(next());
        else if (s == "--cluster-method") cluster_method = next();
        else {
            std::cerr << "unknown argument " << s << std::endl;
            return 2;
        }
    }
    if (cands.empty() || min_ed < 0 || (graph_only && (pair_file.empty() || refine_k))) {
        std::cerr << "usage: igc_cli --cands FILE --min-ed D [--order-threads T] [--out DIR] [--verify] "
                     "[--pair-file PATH] [--refine K ...] | --pair-file PATH --graph-only" << std::endl;
        return 2;
    }
    std::ifstream in(cands);
    if (!in) {
        std::cerr << "cannot open " << cands << std::endl;
        return 1;
    }
    std::vector<uint8_t> syms;
    std::string line;
    uint32_t N = 0, n = 0;
    while (std::getline(in, line)) {
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
        if (line.empty()) continue;
        if (n == 0) n = (uint32_t)line.size();
        if (line.size() != n) {
            std::cerr << "line " << N + 1 << " has length " << line.size() << ", expected " << n << std::endl;
            return 1;
        }
        for (char c : line) {
            const int v = sym_code(c);
            if (v < 0) {
                std::cerr << "bad symbol '" << c << "' on line " << N + 1 << std::endl;
                return 1;
            }
            syms.push_back((uint8_t)v);
        }
        ++N;
    }
    std::cout << "Number of Candidates: " << N << " (length " << n << ")" << std::endl;
    std::cout << "Backend: " << igc_backend() << std::endl;

    igc_ctx *ctx = nullptr;
    if (igc_ctx_create(0, &ctx) != IGC_OK) return die("igc_ctx_create");
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    std::vector<uint32_t> order(N > 0 ? N : 1);
    uint32_t count = 0;
    long long ones = 0;
    double fill_s = 0, proc_s = 0;
    int cluster_iters = -1;
    if (refine_k) {
        igc_refine_stats st;
        const unsigned hw = std::thread::hardware_concurrency();
        if (igc_refine(ctx, syms.data(), N, n, (uint32_t)min_ed, refine_k, km_iters, rounds, converge, 1,
                       hw ? hw : 4, order.data(), &count, &st) != IGC_OK)
            return die("igc_refine");
        for (uint32_t r = 0; r < st.rounds; ++r) {
            std::cout << "round " << r + 1 << ": " << st.round[r].candidates << " candidates, "
                      << st.round[r].clusters << " clusters -> " << st.round[r].codewords << " codewords" << std::endl;
            // the reference's clustered header: clustering time, then solving
            // (graphs + greedy) time (src/SparseMat.cpp:1113-1118)
            fill_s += st.round[r].cluster_s;
            proc_s += st.round[r].graph_s + st.round[r].greedy_s;
        }
        // rounds to convergence; 0 when the round limit ended the loop first
        cluster_iters = st.converged ? (int)st.rounds : 0;
    } else {
        igc_csr g;
        const uint32_t flags = order_t > 1 ? IGC_ORDER_THREADS(order_t) : IGC_ORDER_ASCENDING;
        if (igc_graph_build(ctx, syms.data(), N, n, (uint32_t)min_ed, flags, &g) != IGC_OK)
            return die("igc_graph_build");
        const auto t1 = clk::now();
        fill_s = std::chrono::duration<double>(t1 - t0).count();
        ones = 2 * (long long)g.n_edges;
        std::cout << "Fill AdjList Time:\t" << fill_s << "\tseconds" << std::endl;
        if (!pair_file.empty() && igc_write_pair_file(&g, pair_file.c_str()) != IGC_OK) return die("pair file");
        if (graph_only) {
            std::cout << "Edges: " << g.n_edges << std::endl;
            igc_csr_free(&g);
            igc_ctx_destroy(ctx);
            return 0;
        }
        if (igc_greedy(&g, order.data(), &count) != IGC_OK) return die("igc_greedy");
        proc_s = std::chrono::duration<double>(clk::now() - t1).count();
        std::cout << "Process Matrix Time:\t" << proc_s << "\tseconds" << std::endl;
        igc_csr_free(&g);
    }
    std::cout << "Number of Code Words: " << count << std::endl;
    std::vector<uint8_t> words;
    igc_codebook_info info{};
    info.code_len = (int)n;
    info.min_ed = min_ed;
    info.method_line = "Using Generation Method: FileRead (igc_cli)";
    info.candidate_num = N;
    info.matrix_ones = ones;
    info.threads = order_t;
    info.fill_s = fill_s;
    info.process_s = proc_s;
    info.overall_s = std::chrono::duration<double>(clk::now() - t0).count();
    if (refine_k) {
        info.cluster_k = (int)refine_k;
        info.cluster_iterations = cluster_iters;
        info.convergence_iterations = (int)converge;
        info.cluster_method = cluster_method.c_str();
        info.matrix_ones = 0;
    }
    char name[128];
    if (igc_write_codebook(out.c_str(), syms.data(), n, order.data(), count, &info, name, sizeof name) != IGC_OK)
        return die("igc_write_codebook");
    std::cout << "Wrote " << out << "/" << name << std::endl;
    if (verify) {
        std::vector<uint8_t> cb((size_t)count * n);
        for (uint32_t k = 0; k < count; ++k) std::memcpy(&cb[(size_t)k * n], &syms[(size_t)order[k] * n], n);
        uint64_t e = 0;
        if (igc_graph_count(ctx, cb.data(), count, n, (uint32_t)min_ed, &e) != IGC_OK) return die("verify");
        std::cout << "Vector distance " << (e == 0 ? "SUCCESS" : "FAILURE") << " (" << e << " violations)" << std::endl;
    }
    igc_ctx_destroy(ctx);
    return 0;
}
