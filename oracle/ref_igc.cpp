// ref_igc.cpp -- the reference's own host code calling libigc.so in process:
// the FillAdjListGPU replacement body that INTEGRATION.md gives a maintainer
// (it replaces src/SparseMat.cpp:454-534), compiled against the reference's
// AdjList and run through its RowsBySum and Codebook (src/SparseMat.cpp:81-284,
// 679-731), exactly as CodebookAdjList (:734-777) sequences them.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ (needs
// the reference tree and a built indexgen_amd/libigc.so); a GPU test compares
// its codebook with the reference's CPU path (ref_driver graph 1 ...).
//
//   ref_igc cands minED dir    prints "ones <2|E|> fill_s <t> greedy_s <t>",
//                              writes dir/codebook.txt (selection order)
#include "IndexGen.hpp"
#include "SparseMat.hpp"
#include "igc.h"

#include <chrono>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <mutex>
#include <sstream>
#include <string>
#include <unistd.h>
#include <vector>

using namespace std;

void Codebook(AdjList &adjList, vector<string> &codebook, const vector<string> &candidates, const int saveInterval,
              const bool resume);

// ---- the body INTEGRATION.md shows (FillAdjListGPU's signature, one library call) ----
void FillAdjListGPU_igc(AdjList &adjList, const vector<string> &candidates, const int minED,
                        long long int &matrixOnesNum, const string & = "", double = 10.0, bool silent = false,
                        bool isBinary = false)
{
    static igc_ctx *ctx = nullptr; // one context per process (thread-safe calls)
    static std::once_flag once;
    std::call_once(once, [] {
        if (igc_ctx_create(0, &ctx) != IGC_OK) {
            std::cerr << igc_last_error() << endl;
            exit(1);
        }
    });
    // binary candidates ('0'/'1') of any length up to 64 and quaternary ones up to 32 use
    // the same call: the library packs n > 32 as binary words and refuses other symbols
    const uint32_t N = candidates.size(), n = N ? candidates[0].size() : 1;
    if (!isBinary && n > IGC_MAX_QUATERNARY_LEN) {
        std::cerr << "L > 32" << endl;
        exit(1);
    }
    vector<uint8_t> syms((size_t)N * n);
    for (uint32_t i = 0; i < N; ++i)
        for (uint32_t p = 0; p < n; ++p) syms[(size_t)i * n + p] = candidates[i][p] - '0';
    igc_csr g;
    // IGC_ORDER_THREADS(T) reproduces the neighbour order of FillAdjList with T threads;
    // 0 (ascending) equals the CPU path with one thread.
    if (igc_graph_build(ctx, syms.data(), N, n, minED, 0, &g) != IGC_OK) {
        std::cerr << "GPU graph build failed: " << igc_last_error() << endl;
        exit(1); // the reference's behaviour on failure
    }
    adjList.Init(N);
    for (uint32_t x = 0; x < N; ++x)
        for (uint64_t q = g.row_ptr[x]; q < g.row_ptr[x + 1]; ++q) adjList.Set(x, g.col[q]);
    matrixOnesNum = 2 * (long long)g.n_edges;
    igc_csr_free(&g);
    (void)silent;
}
// ---- end of the INTEGRATION.md body ----

static vector<string> read_lines(const string &path)
{
    ifstream in(path);
    vector<string> out;
    string s;
    while (getline(in, s)) {
        if (!s.empty() && s.back() == '\r') s.pop_back();
        if (!s.empty()) out.push_back(s);
    }
    return out;
}

int main(int argc, char **argv)
{
    if (argc != 4) {
        std::cerr << "usage: ref_igc cands minED dir" << endl;
        return 2;
    }
    const vector<string> cands = read_lines(argv[1]);
    const int minED = std::atoi(argv[2]);
    if (chdir(argv[3]) != 0) {
        std::cerr << "chdir failed" << endl;
        return 1;
    }
    AdjList adj;
    long long ones = 0;
    auto t0 = chrono::steady_clock::now();
    FillAdjListGPU_igc(adj, cands, minED, ones);
    adj.RowsBySum(); // CodebookAdjList's next step (src/SparseMat.cpp:759)
    auto t1 = chrono::steady_clock::now();
    vector<string> codebook;
    streambuf *old = std::cout.rdbuf();
    ostringstream sink;
    std::cout.rdbuf(sink.rdbuf());
    Codebook(adj, codebook, cands, 0, false);
    auto t2 = chrono::steady_clock::now();
    std::cout.rdbuf(old);
    ofstream cb("codebook.txt");
    for (const auto &w : codebook) cb << w << '\n';
    std::cout << "ones " << ones << " fill_s " << chrono::duration<double>(t1 - t0).count() << " greedy_s "
              << chrono::duration<double>(t2 - t1).count() << endl;
    return 0;
}
