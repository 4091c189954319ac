// ref_driver.cpp -- a small command-line front end, written for this project,
// over the IndexGen reference's own C++ sources (compiled in place from the
// reference tree by oracle/Makefile into oracle/_ref/).
//
// TEST INFRASTRUCTURE ONLY: it produces golden vectors (tests/golden) and the
// CPU baseline's "reference" timing.  Nothing in the shipped library uses it.
//
// Modes
//   pairs                      stdin: lines "k text pattern" (symbols '0'..'3')
//                              stdout: "banded wf" per line, where banded is
//                              EditDistanceBanded(text, MakePattern(pattern), k)
//                              (include/EditDistance.hpp:327-330) and wf is
//                              FastEditDistance(text, pattern) (src/Utils.cpp:15-51)
//   graph T minED cands dir    FillAdjList with T threads (src/SparseMat.cpp:567-615),
//                              then AdjList::ToFile -> dir/adj.txt (neighbour
//                              lists in list order), Codebook (:679-731) run
//                              inside dir -> dir/codebook.txt (selection order);
//                              prints "ones <2|E|> fill_s <t> greedy_s <t>"
//   fill T minED cands         FillAdjList only, prints "ones <n> fill_s <t>"
//                              (used as the timed CPU baseline)
//   digest T minED cands dir   FillAdjList with T threads, then the edge set as
//                              dir/upper.bin (LE uint32 pairs (i, j), i < j,
//                              sorted by (i, j); read back from AdjList::ToFile)
//                              and Codebook run inside dir -> dir/codebook.txt;
//                              prints "ones <2|E|> fill_s <t> greedy_s <t>".
//                              For full-size sets, where adj.txt would be large
//                              to keep as a fixture.
//   solve T minED cands        the reference's CPU codebook path without files:
//                              FillAdjList (T threads), then Codebook's
//                              selection loop (AdjList::FindMinDel until the
//                              list is empty, then the remaining nodes,
//                              src/SparseMat.cpp:698-729) without its progress
//                              dumps; prints "ones <2|E|> fill_s <t> greedy_s <t>
//                              codewords <c>" (the CPU baseline of the bench)
//   gpu minED cands dir        the reference's own CodebookAdjList (src/SparseMat.cpp:734-777)
//                              with useGPU = true, run inside dir: its
//                              FillAdjListGPU writes the candidates, runs
//                              $INDEXGEN_PYTHON $INDEXGEN_ROOT/src/gpu_graph_generator.py
//                              and reads the pair file back (:454-534); with
//                              INDEXGEN_ROOT pointing at a tree whose script is
//                              indexgen_amd/gpu_graph_generator.py this drives
//                              libigc.so through the reference's unmodified
//                              host.  Writes dir/codebook.txt; prints "ones <n>
//                              fill_s <t> greedy_s <t>"
//   header words n minED maxRun minGC maxGC T candfile candNum ones K iters conv method cand_s fill_s proc_s all_s dir
//                              the reference's own codebook file: ToFile
//                              (src/Utils.cpp:394-418) of the words in `words`
//                              (one per line) with Params for the FileRead
//                              method (file=candfile), clustering.method /
//                              convergenceIterations = method / conv, clusterK
//                              = K and clusterIterations = iters (K <= 0: an
//                              unclustered run), written inside dir
//   linear n minHD bias rowp colp maxRun minGC maxGC
//                              Candidates(params) for the LinearCode generator
//                              with MANUAL vectors (src/Candidates.cpp:178-188),
//                              one candidate per line
//   vt n a b T maxRun minGC maxGC
//   dvt n s T maxRun minGC maxGC
//   all n maxRun minGC maxGC   the same for the VTCode, DifferentialVTCode and
//                              AllStrings generators (T threads)
#include "CandidateGenerator.hpp"
#include "Candidates.hpp"
#include "EditDistance.hpp"
#include "IndexGen.hpp"
#include "SparseMat.hpp"
#include "Utils.hpp"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <unistd.h>
#include <vector>

// Defined with external linkage in src/SparseMat.cpp but not declared in its header.
void FillAdjList(AdjList &adjList, const std::vector<std::string> &candidates, const int minED, const int threadNum,
                 const int saveInterval, const bool resume, long long int &matrixOnesNum);
void Codebook(AdjList &adjList, std::vector<std::string> &codebook, const std::vector<std::string> &candidates,
              const int saveInterval, const bool resume);
std::string FileName(const int codeLen, const int codeSize, const int minED); // src/Utils.cpp:381-392
void CodebookAdjList(const std::vector<std::string> &candidates, std::vector<std::string> &codebook, const int minED,
                     const int threadNum, const int saveInterval, long long int &matrixOnesNum,
                     std::chrono::duration<double> &fillAdjListTime, std::chrono::duration<double> &processMatrixTime,
                     const bool useGPU, double maxGPUMemoryGB, const std::string &candFilename, bool isBinary);

static std::vector<std::string> read_lines(const std::string &path)
{
    std::ifstream in(path);
    std::vector<std::string> out;
    std::string s;
    while (std::getline(in, s)) {
        if (!s.empty() && s.back() == '\r') s.pop_back();
        if (!s.empty()) out.push_back(s);
    }
    return out;
}

static std::vector<int> parse_csv(const std::string &s)
{
    std::vector<int> v;
    if (s == "-") return v;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) v.push_back(std::stoi(tok));
    return v;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: ref_driver pairs|graph|fill|linear ..." << std::endl;
        return 2;
    }
    std::string mode = argv[1];
    if (mode == "pairs") {
        int k;
        std::string text, pat;
        std::ios::sync_with_stdio(false);
        while (std::cin >> k >> text >> pat) {
            PatternHandle H = MakePattern(pat);
            int banded = EditDistanceBanded(text, H, k);
            int wf = FastEditDistance(text, pat);
            std::cout << banded << ' ' << wf << '\n';
        }
        return 0;
    }
    if (mode == "digest" && argc == 6) {
        int T = std::atoi(argv[2]);
        int minED = std::atoi(argv[3]);
        std::vector<std::string> cands = read_lines(argv[4]);
        std::string dir = argv[5];
        AdjList adj;
        long long ones = 0;
        auto t0 = std::chrono::steady_clock::now();
        FillAdjList(adj, cands, minED, T, 0, false, ones);
        auto t1 = std::chrono::steady_clock::now();
        const double fill_s = std::chrono::duration<double>(t1 - t0).count();
        {
            // the adjacency as the reference itself writes it ("i\tj" per list entry)
            const std::string adj_path = dir + "/adj.txt";
            adj.ToFile(adj_path);
            FILE *f = std::fopen(adj_path.c_str(), "r");
            if (!f) {
                std::cerr << "cannot read " << adj_path << std::endl;
                return 1;
            }
            std::vector<uint64_t> keys;
            keys.reserve((size_t)(ones / 2));
            long long a, b;
            long long entries = 0;
            while (std::fscanf(f, "%lld %lld", &a, &b) == 2) {
                ++entries;
                if (a < b) keys.push_back(((uint64_t)a << 32) | (uint64_t)b);
            }
            std::fclose(f);
            std::remove(adj_path.c_str());
            if (entries != ones || (long long)keys.size() * 2 != ones) {
                std::cerr << "adjacency file holds " << entries << " entries, " << keys.size()
                          << " with i < j; FillAdjList reported " << ones << std::endl;
                return 1;
            }
            std::sort(keys.begin(), keys.end());
            FILE *o = std::fopen((dir + "/upper.bin").c_str(), "wb");
            if (!o) {
                std::cerr << "cannot write upper.bin" << std::endl;
                return 1;
            }
            std::vector<uint32_t> buf;
            buf.reserve(1 << 20);
            for (uint64_t k : keys) {
                buf.push_back((uint32_t)(k >> 32));
                buf.push_back((uint32_t)k);
                if (buf.size() >= (1u << 20)) {
                    std::fwrite(buf.data(), 4, buf.size(), o);
                    buf.clear();
                }
            }
            std::fwrite(buf.data(), 4, buf.size(), o);
            std::fclose(o);
        }
        if (chdir(dir.c_str()) != 0) {
            std::cerr << "chdir failed" << std::endl;
            return 1;
        }
        std::vector<std::string> codebook;
        std::streambuf *old = std::cout.rdbuf();
        std::ostringstream sink;
        std::cout.rdbuf(sink.rdbuf());
        auto t2 = std::chrono::steady_clock::now();
        Codebook(adj, codebook, cands, 0, false);
        auto t3 = std::chrono::steady_clock::now();
        std::cout.rdbuf(old);
        std::ofstream cb("codebook.txt");
        for (const auto &w : codebook) cb << w << '\n';
        std::cout << "ones " << ones << " fill_s " << fill_s << " greedy_s "
                  << std::chrono::duration<double>(t3 - t2).count() << std::endl;
        return 0;
    }
    if (mode == "gpu" && argc == 5) {
        const int minED = std::atoi(argv[2]);
        std::vector<std::string> cands = read_lines(argv[3]);
        if (chdir(argv[4]) != 0) {
            std::cerr << "chdir failed" << std::endl;
            return 1;
        }
        std::vector<std::string> codebook;
        long long ones = 0;
        std::chrono::duration<double> fill{}, proc{};
        std::streambuf *old = std::cout.rdbuf();
        std::ostringstream sink;
        std::cout.rdbuf(sink.rdbuf());
        CodebookAdjList(cands, codebook, minED, 1, 0, ones, fill, proc, true, 10.0, "", false);
        std::cout.rdbuf(old);
        std::ofstream cb("codebook.txt");
        for (const auto &w : codebook) cb << w << '\n';
        std::cout << "ones " << ones << " fill_s " << fill.count() << " greedy_s " << proc.count() << std::endl;
        return 0;
    }
    if (mode == "solve" && argc == 5) {
        const int T = std::atoi(argv[2]);
        const int minED = std::atoi(argv[3]);
        std::vector<std::string> cands = read_lines(argv[4]);
        AdjList adj;
        long long ones = 0;
        auto t0 = std::chrono::steady_clock::now();
        FillAdjList(adj, cands, minED, T, 0, false, ones); // ends with RowsBySum
        auto t1 = std::chrono::steady_clock::now();
        std::vector<bool> remaining(cands.size(), true);
        double t_min = 0, t_del = 0;
        size_t words = 0;
        while (!adj.empty()) {
            (void)adj.FindMinDel(remaining, t_min, t_del);
            ++words;
        }
        for (bool r : remaining) words += r ? 1 : 0;
        auto t2 = std::chrono::steady_clock::now();
        std::cout << "ones " << ones << " fill_s " << std::chrono::duration<double>(t1 - t0).count() << " greedy_s "
                  << std::chrono::duration<double>(t2 - t1).count() << " codewords " << words << std::endl;
        return 0;
    }
    if (mode == "header" && argc == 21) {
        std::vector<std::string> words = read_lines(argv[2]);
        const int n = std::atoi(argv[3]);
        Params params(n, std::atoi(argv[4]), std::atoi(argv[5]), std::atof(argv[6]), std::atof(argv[7]),
                      std::atoi(argv[8]), 0, GenerationMethod::FILE_READ,
                      std::make_unique<FileReadConstraints>(std::string(argv[9])));
        const int K = std::atoi(argv[12]);
        params.clustering.enabled = K > 0;
        params.clustering.k = K;
        params.clustering.convergenceIterations = std::atoi(argv[14]);
        params.clustering.method = argv[15];
        using dur = std::chrono::duration<double>;
        if (chdir(argv[20]) != 0) {
            std::cerr << "chdir failed" << std::endl;
            return 1;
        }
        ToFile(words, params, std::atoi(argv[10]), std::atoll(argv[11]), dur(std::atof(argv[16])),
               dur(std::atof(argv[17])), dur(std::atof(argv[18])), dur(std::atof(argv[19])), K, std::atoi(argv[13]));
        std::cout << FileName(n, (int)words.size(), std::atoi(argv[4])) << std::endl;
        return 0;
    }
    if ((mode == "graph" && argc == 6) || (mode == "fill" && argc == 5)) {
        int T = std::atoi(argv[2]);
        int minED = std::atoi(argv[3]);
        std::vector<std::string> cands = read_lines(argv[4]);
        AdjList adj;
        long long ones = 0;
        auto t0 = std::chrono::steady_clock::now();
        // FillAdjList prints nothing; it writes/deletes per-thread progress files
        // only when saveInterval > 0.
        FillAdjList(adj, cands, minED, T, 0, false, ones);
        auto t1 = std::chrono::steady_clock::now();
        double fill_s = std::chrono::duration<double>(t1 - t0).count();
        if (mode == "fill") {
            std::cout << "ones " << ones << " fill_s " << fill_s << std::endl;
            return 0;
        }
        std::string dir = argv[5];
        adj.RowsBySum();
        adj.ToFile(dir + "/adj.txt");
        // Codebook() writes and removes progress_* files in the working directory.
        if (chdir(dir.c_str()) != 0) {
            std::cerr << "chdir failed" << std::endl;
            return 1;
        }
        std::vector<std::string> codebook;
        std::streambuf *old = std::cout.rdbuf();
        std::ostringstream sink;
        std::cout.rdbuf(sink.rdbuf());
        auto t2 = std::chrono::steady_clock::now();
        Codebook(adj, codebook, cands, 0, false);
        auto t3 = std::chrono::steady_clock::now();
        std::cout.rdbuf(old);
        std::ofstream cb("codebook.txt");
        for (const auto &w : codebook) cb << w << '\n';
        std::cout << "ones " << ones << " fill_s " << fill_s << " greedy_s "
                  << std::chrono::duration<double>(t3 - t2).count() << std::endl;
        return 0;
    }
    if (mode == "linear" && argc == 10) {
        int n = std::atoi(argv[2]);
        int minHD = std::atoi(argv[3]);
        auto lc = std::make_unique<LinearCodeConstraints>(minHD);
        lc->bias = parse_csv(argv[4]);
        lc->row_perm = parse_csv(argv[5]);
        lc->col_perm = parse_csv(argv[6]);
        lc->bias_mode = lc->bias.empty() ? VectorMode::DEFAULT : VectorMode::MANUAL;
        lc->row_perm_mode = lc->row_perm.empty() ? VectorMode::DEFAULT : VectorMode::MANUAL;
        lc->col_perm_mode = lc->col_perm.empty() ? VectorMode::DEFAULT : VectorMode::MANUAL;
        Params params(n, 3, std::atoi(argv[7]), std::atof(argv[8]), std::atof(argv[9]), 1, 0,
                      GenerationMethod::LINEAR_CODE, std::move(lc));
        std::streambuf *old = std::cout.rdbuf();
        std::ostringstream sink;
        std::cout.rdbuf(sink.rdbuf());
        std::vector<std::string> cands = Candidates(params);
        std::cout.rdbuf(old);
        for (const auto &s : cands) std::cout << s << '\n';
        return 0;
    }
    if ((mode == "vt" && argc == 9) || (mode == "dvt" && argc == 8) || (mode == "all" && argc == 6)) {
        const int n = std::atoi(argv[2]);
        std::unique_ptr<GenerationConstraints> c;
        GenerationMethod m;
        int T = 1, f = 3;
        if (mode == "vt") {
            c = std::make_unique<VTCodeConstraints>(std::atoi(argv[3]), std::atoi(argv[4]));
            m = GenerationMethod::VT_CODE;
            T = std::atoi(argv[5]);
            f = 6;
        } else if (mode == "dvt") {
            c = std::make_unique<DifferentialVTCodeConstraints>(std::atoi(argv[3]));
            m = GenerationMethod::DIFFERENTIAL_VT_CODE;
            T = std::atoi(argv[4]);
            f = 5;
        } else {
            c = std::make_unique<AllStringsConstraints>();
            m = GenerationMethod::ALL_STRINGS;
        }
        Params params(n, 3, std::atoi(argv[f]), std::atof(argv[f + 1]), std::atof(argv[f + 2]), T, 0, m, std::move(c));
        std::streambuf *old = std::cout.rdbuf();
        std::ostringstream sink;
        std::cout.rdbuf(sink.rdbuf());
        std::vector<std::string> cands = Candidates(params);
        std::cout.rdbuf(old);
        for (const auto &x : cands) std::cout << x << '\n';
        return 0;
    }
    std::cerr << "bad arguments" << std::endl;
    return 2;
}
