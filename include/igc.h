/*
 * igc.h -- C ABI of the MI355X conflict-graph engine (libigc.so).
 *
 * The engine builds IndexGen's edit-distance conflict graph on a gfx950 GPU:
 * for equal-length candidate strings over {A,C,G,T} (symbol codes 0..3) it
 * adds edge (i,j) iff Levenshtein(cand[i], cand[j]) < edit_dist, exactly the
 * predicate of the reference's EditDistanceBandedAtLeast
 * (include/EditDistance.hpp:368-372, used at src/SparseMat.cpp:544-554).
 *
 * It replaces the reference's GPU hand-off, which is a process and file
 * boundary:
 *   FillAdjListGPU(AdjList&, const vector<string>& candidates, int minED,
 *                  long long& matrixOnesNum, const string& inputFilename,
 *                  double maxGPUMemoryGB, bool silent, bool isBinary)
 *   (src/SparseMat.cpp:454-534) -> system("python3 src/gpu_graph_generator.py
 *   cands.txt edges.bin minED maxGB") -> LE int32 pair file -> FromBinaryFile
 *   (src/SparseMat.cpp:338-390).
 * Here the same inputs (candidate list, n, editDist) go in by pointer and the
 * same output (the symmetric adjacency, matrixOnesNum = 2*|E|) comes back as a
 * CSR, with neighbour lists in a declared canonical order.
 *
 * Conventions
 *   - Every call returns an int status (IGC_OK = 0).  The library never calls
 *     exit() or abort(); igc_last_error() gives a message (thread-local).
 *   - Inputs are caller-owned and only read during the call.
 *   - Host CSR outputs are library-allocated and released with igc_csr_free().
 *   - A context (igc_ctx) owns one device, one HIP stream and cached device
 *     buffers; calls on one context are serialised by an internal mutex, calls
 *     on different contexts may run concurrently from different host threads.
 *   - Neighbour order: ascending by default, which is the order the reference
 *     CPU path produces with one thread (src/SparseMat.cpp:536-615, T = 1).
 *     IGC_ORDER_THREADS(T) reproduces the order of its T-thread merge
 *     (key (min(x,y) mod T, min(x,y), max(x,y)) within row x).  The greedy
 *     codebook depends on this order (src/SparseMat.cpp:122-127, 207-251).
 */
#ifndef IGC_H_
#define IGC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IGC_API __attribute__((visibility("default")))

/* ---- status codes ---- */
#define IGC_OK 0
#define IGC_ERR_INVALID 1   /* bad argument (size, symbol > 3, n out of range ...) */
#define IGC_ERR_OOM 2       /* device or host allocation failed */
#define IGC_ERR_DEVICE 3    /* HIP runtime error or no usable gfx950 device */
#define IGC_ERR_ADMISSION 4 /* the job does not fit the configured HBM budget */
#define IGC_ERR_IO 5        /* file could not be written/read */
#define IGC_ERR_BUSY 6      /* result not ready yet (server poll) */

/* longest candidate the kernels accept: the reference's binary GPU solver's
 * limit (L <= 64, src/cuda_edit_distance_binary.py:272).  Strings up to
 * IGC_MAX_QUATERNARY_LEN take symbols 0..3 (the quaternary solver's limit,
 * L <= 32, src/cuda_edit_distance.py:287) and are packed as two symbol bit
 * planes, in a 32-bit word for n <= 16 and a 64-bit word above; longer strings
 * must be binary (symbols 0/1, the reference's isBinary path) and are packed
 * one bit per symbol in a 64-bit word.  Binary strings of length <= 32 take the
 * quaternary path, which gives the same distances. */
#define IGC_MAX_LEN 64
#define IGC_MAX_QUATERNARY_LEN 32

/* ---- flags for igc_graph_build ---- */
#define IGC_ORDER_ASCENDING 0u
/* neighbour order of the reference's FillAdjList with T threads, 1 <= T <= 65535 */
#define IGC_ORDER_THREADS(T) (((uint32_t)(T)&0xFFFFu) << 16)
/* CSR of the upper triangle only: row i lists its neighbours j > i, ascending
 * (|E| entries instead of 2|E|; the pair-file order of the reference) */
#define IGC_UPPER_TRIANGLE 1u
/* igc_graph_build: leave the CSR in device memory owned by the context (as
 * igc_graph_build_device does) instead of copying it to the host; the
 * symbols still come from host memory (pinned memory makes the copy an
 * asynchronous DMA).  Ascending order only. */
#define IGC_RESULT_ON_DEVICE 2u

/* where a CSR's arrays live */
#define IGC_ON_HOST 0u
#define IGC_ON_DEVICE 1u

typedef struct igc_ctx igc_ctx;

/* Symmetric adjacency of the conflict graph.  n_edges counts unordered pairs;
 * the reference's matrixOnesNum ("Number of Ones in Matrix") is 2*n_edges. */
typedef struct igc_csr {
    uint32_t n_nodes;
    uint32_t location; /* IGC_ON_HOST or IGC_ON_DEVICE */
    uint64_t n_edges;
    uint64_t *row_ptr; /* n_nodes + 1 entries */
    uint32_t *col;     /* 2 * n_edges entries */
} igc_csr;

/* Timings of the last build on a context, from HIP events on its stream. */
typedef struct igc_timing {
    float pack_ms;      /* symbol packing kernel */
    float pairs_ms;     /* all-pairs edit-distance kernel (the hot kernel) */
    float csr_ms;       /* degree + scan + scatter + row sort */
    float total_ms;     /* first event to last event of the build */
    uint64_t pairs;     /* N(N-1)/2 pairs tested */
    uint64_t n_edges;
    uint32_t launches;  /* kernels launched by the last build */
    uint32_t retries;   /* edge-buffer regrowths in the last build */
} igc_timing;

/* ---- context ---- */

/* Creates a context on HIP device `device`.  Fails with IGC_ERR_DEVICE when no
 * gfx950 device is present (no CPU fallback exists). */
IGC_API int igc_ctx_create(int device, igc_ctx **out);
IGC_API int igc_ctx_destroy(igc_ctx *ctx);
/* Use an existing hipStream_t (as void*) instead of the context's own; NULL
 * restores the context's own stream. */
IGC_API int igc_ctx_set_stream(igc_ctx *ctx, void *hip_stream);
IGC_API void *igc_ctx_stream(igc_ctx *ctx);
/* Exact row skipping (single sets and batched sets): candidates are ordered by
 * the symbol counts of their two halves, and a row is skipped for a wave's
 * column window when those counts prove no column of the window is within the
 * threshold.  The graph is the same either way.  IGC_PREFILTER_AUTO (default)
 * sorts sets of at least IGC_PREFILTER_MIN_N strings, where the rows saved
 * outweigh the sort; IGC_PREFILTER_ALWAYS sorts every set with a nontrivial
 * threshold.  No counterpart in the reference (an internal speed-up of
 * FillAdjListGPU). */
#define IGC_PREFILTER_OFF 0
#define IGC_PREFILTER_AUTO 1
#define IGC_PREFILTER_ALWAYS 2
#define IGC_PREFILTER_MIN_N 16384u
IGC_API int igc_ctx_set_prefilter(igc_ctx *ctx, int mode);
/* Pair predicate of the context's kernels: IGC_PRED_AUTO (default) picks the
 * cheaper of Landau-Vishkin and Myers for (n, editDist); IGC_PRED_LANDAU_VISHKIN
 * forces Landau-Vishkin wherever it is instantiated (editDist - 1 <= 13) and
 * IGC_PRED_MYERS forces Myers; editDist > n always takes every pair.  The graph
 * is the same either way (exact predicates); for measurement and tests.  No
 * counterpart in the reference. */
#define IGC_PRED_AUTO (-1)
IGC_API int igc_ctx_set_predicate(igc_ctx *ctx, int predicate);
/* How the CSR is assembled (same CSR either way).  IGC_CSR_AUTO (default): a
 * split pass on the row's high bits and 32-bit radix passes when the row's
 * low bits and the neighbour fit 32 bits with a split of at most 9 bits
 * (single sets up to 2^20 strings; batches up to 2^32 / largest set nodes),
 * else 64-bit keys; IGC_CSR_WIDE_KEYS: always 64-bit keys.  No counterpart in
 * the reference (its host AdjList is filled pair by pair). */
#define IGC_CSR_AUTO 0
#define IGC_CSR_WIDE_KEYS 1
IGC_API int igc_ctx_set_csr_form(igc_ctx *ctx, int form);
/* Capacity of a build's first edge buffer, in edges (0, the default: twice
 * the expected edge count of random strings of that length and threshold, or
 * 2^26 edges where no estimate exists).  A build whose graph has more edges
 * regrows the buffer to the exact count and runs its pair kernels again, so
 * the result never depends on it (igc_timing.retries counts the regrowths).
 * For tests of that path and for callers that know their graph's density.
 * No counterpart in the reference (its per-tile edge buffer is fixed,
 * src/cuda_edit_distance.py:320). */
IGC_API int igc_ctx_set_edge_cap(igc_ctx *ctx, uint64_t edges);
/* igc_refine replays one captured hipGraph per round for its device work up
 * to the cluster assignment (on, the default), or launches the same kernels
 * directly (off); the results are the same.  For tests and measurement; no
 * counterpart in the reference. */
IGC_API int igc_ctx_set_graph_replay(igc_ctx *ctx, int on);
IGC_API int igc_last_timing(igc_ctx *ctx, igc_timing *out);

/* ---- single candidate set ---- */

/* Builds the conflict graph of N candidates of length n (1 <= n <= 64) given
 * as an N x n row-major array of symbol codes in host memory: 0..3 for
 * n <= 32, 0/1 for n > 32 (IGC_ERR_INVALID otherwise).  edit_dist is the
 * reference's minED: edge iff Levenshtein < edit_dist.  The CSR comes back in
 * host memory.
 * Replaces FillAdjListGPU (src/SparseMat.cpp:454-534) and the Python solvers
 * it runs (src/gpu_graph_generator.py:17-80, src/cuda_edit_distance.py:278-447;
 * with isBinary, src/gpu_graph_generator_binary.py:17-54,
 * src/cuda_edit_distance_binary.py:285-389). */
IGC_API int igc_graph_build(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                            uint32_t flags, igc_csr *out);

/* Same, with d_syms already in device memory and the CSR left in device
 * memory owned by the context (valid until the next build on it; do not free).
 * Ascending order only.  Runs asynchronously on the context's stream except
 * for one final read of the edge count. */
IGC_API int igc_graph_build_device(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n,
                                   uint32_t edit_dist, igc_csr *out);

/* Same as igc_graph_build_device, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): no counterpart in the reference.  The CSR goes to caller device
 * buffers -- d_row_ptr (N + 1 entries) and d_col (col_cap entries, >= 2|E|) --
 * and every temporary comes from a fixed device array of 128 MiB that each GPU
 * the library runs on holds until the process ends; the call allocates nothing,
 * host or device.  Limit: packed words (4 or 8 B per string), edge list (8 B
 * per edge) and sort temporaries (about 16 B per edge) within the 128 MiB,
 * about 5 million edges; beyond it, or when col_cap < 2|E|, IGC_ERR_INVALID
 * with the size needed (*n_edges is set once the edges are counted).
 * Ascending order, no row filter (same graph).  Calls on one GPU share the
 * array and are serialised by the library; the array keeps the last call's
 * contents (what must start at zero is cleared by a kernel each call). */
IGC_API int igc_graph_build_device_noalloc(igc_ctx *ctx, const uint8_t *d_syms, uint32_t N, uint32_t n,
                                           uint32_t edit_dist, uint64_t *d_row_ptr, uint32_t *d_col,
                                           uint64_t col_cap, uint64_t *n_edges);

/* One timed tile shape of the pair kernel: `cpl` register slots per lane
 * (1, 2, 4; each slot tests lane_cols columns: 2 for n <= 16, 1 above), tiles
 * of 256*cpl*lane_cols columns and 1/q as many rows, best time over the
 * repetitions. */
typedef struct igc_tune_entry {
    int cpl;
    uint32_t q;
    float ms;
    float gpairs_s;
} igc_tune_entry;

/* Autotuner: times every tile shape of the pair kernel on N random strings of
 * length n at edit_dist (device-generated, count-only kernel, best of `reps`
 * after a warm-up) and makes the fastest the context's shape for that
 * threshold and size class (sizes bucketed by powers of 4 from 4096).  Writes
 * up to `cap` entries and the number tried to *count. */
IGC_API int igc_autotune(igc_ctx *ctx, uint32_t n, uint32_t edit_dist, uint32_t N, uint32_t reps,
                         igc_tune_entry *entries, uint32_t cap, uint32_t *count);
/* Shape the context will use for a set of N strings at (n, edit_dist). */
IGC_API int igc_tuned_shape(igc_ctx *ctx, uint32_t n, uint32_t edit_dist, uint32_t N, int *cpl, uint32_t *q);

/* The single-set pair launch the context would make for N strings at
 * (n, edit_dist): which test, how many pairs per register, the tile shape and
 * the workgroup count.  Introspection for benchmarks and profiling; no
 * counterpart in the reference. */
#define IGC_PRED_LANDAU_VISHKIN 0
#define IGC_PRED_MYERS 1
#define IGC_PRED_ALL_PAIRS 2
typedef struct igc_pair_plan {
    int predicate;      /* IGC_PRED_* */
    int k;              /* Levenshtein threshold edit_dist - 1 */
    int lane_cols;      /* pairs per register: 2 (16-bit lanes, n <= 16) or 1 */
    int slots;          /* registers of columns per lane */
    uint32_t q;         /* tile_rows = tile_cols / q */
    uint32_t tile_cols; /* 4 waves x 64 lanes x slots x lane_cols */
    uint32_t tile_rows;
    uint64_t tiles;     /* workgroups of the launch (upper triangle) */
    int row_filter;     /* 1: candidates sorted for exact row skipping (slots is then 1) */
    int pad;
} igc_pair_plan;
IGC_API int igc_pair_plan_of(igc_ctx *ctx, uint32_t n, uint32_t edit_dist, uint32_t N, igc_pair_plan *out);

/* Counts the edges without materialising them: the verify mode of the
 * reference (VerifyDist, src/Utils.cpp:203-309: zero edges == SUCCESS). */
IGC_API int igc_graph_count(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                            uint64_t *n_edges);

/* Copies a device CSR (from igc_graph_build_device) to a new host CSR. */
IGC_API int igc_csr_to_host(igc_ctx *ctx, const igc_csr *dev, uint32_t flags, igc_csr *host);

/* Reorders the neighbour lists of an ascending host CSR in place to the order
 * `flags` names (IGC_ORDER_THREADS(T): the reference's T-thread merge order,
 * src/SparseMat.cpp:599-609).  IGC_ORDER_ASCENDING leaves it unchanged. */
IGC_API int igc_csr_reorder(igc_csr *csr, uint32_t flags);

/* Releases a host CSR's arrays (no-op for device CSRs owned by a context).
 * Arrays of 4 MiB and more go back to a process-wide cache of host blocks whose
 * pages are already faulted in (at most min(16 GiB, physical memory / 8)), from
 * which later host CSRs are allocated. */
IGC_API int igc_csr_free(igc_csr *csr);

/* ---- many candidate sets in one launch ---- */

/* One candidate set of a batch: N strings of length n (symbol codes 0..3,
 * row-major), edge iff Levenshtein < edit_dist; flags as for igc_graph_build. */
typedef struct igc_set {
    const uint8_t *syms;
    uint32_t N;
    uint32_t n;
    uint32_t edit_dist;
    uint32_t flags;
} igc_set;

/* Builds the conflict graphs of `count` independent, unevenly sized sets (each
 * with its own n and edit_dist) in one pair launch per predicate (sets of the
 * same predicate share a launch) plus one CSR assembly over their disjoint
 * union; outs[s] is set s's host CSR, identical
 * to what igc_graph_build returns for that set alone.  Replaces one
 * FillAdjListGPU process per cluster in the reference's clustered loop
 * (SolveIndependentSet, src/SparseMat.cpp:784-851, called at :985-1020). */
IGC_API int igc_graph_build_batch(igc_ctx *ctx, const igc_set *sets, uint32_t count, igc_csr *outs);

/* Same with set symbols in device memory; the union CSR (global node ids) stays
 * in device memory owned by the context.  Every non-empty set's `syms` must be
 * device memory of the context's device: the sets are gathered by a kernel that
 * reads them directly (pointers are not validated, to keep the call free of a
 * per-set runtime query).  node_off[count+1] receives each set's
 * first global id (set s owns rows node_off[s]..node_off[s+1]-1, ids shifted
 * by node_off[s]); set_edges[count] its edge count. */
IGC_API int igc_graph_build_batch_device(igc_ctx *ctx, const igc_set *sets, uint32_t count, igc_csr *union_out,
                                         uint32_t *node_off, uint64_t *set_edges);

/* ---- candidate generators on the device ---- */

/* applyFilters (src/CandidateGenerator.cpp:24-50): keep strings whose longest
 * run of one symbol is at most max_run (0: no run filter) and whose fraction of
 * symbols 1 and 2 (C, G) lies in [min_gc, max_gc], compared in double as the
 * reference does (both 0: no GC filter). */
typedef struct igc_filter {
    uint32_t max_run;
    uint32_t pad;
    double min_gc;
    double max_gc;
} igc_filter;

/* Generated candidates: count x n symbols in device memory owned by the
 * context, in the reference generator's order; valid until the next generator
 * call on the context.  Feed them to igc_graph_build_device (count < 2^31). */
typedef struct igc_candidates {
    const uint8_t *d_syms;
    uint64_t count;      /* strings kept */
    uint64_t enumerated; /* words tried */
    uint64_t serial;     /* the context's generation number (checked by igc_candidates_to_host) */
    uint32_t n;
    float ms;            /* device time of the generation (HIP events) */
} igc_candidates;

/* longest word the enumerating generators (VT, differential VT, all strings)
 * accept: 4^21 words */
#define IGC_GEN_MAX_WORD_LEN 21

/* LinearCode candidates: every message m of k symbols in NextBase4 order
 * (src/Utils.cpp:525-545), codeword m*G + bias over GF(4) (labels 0, 1, x,
 * x+1 of GF(2)[x]/(x^2+x+1), src/Candidates/GF4.cpp:15-30), then the filter
 * (NULL: none).  `gen` is the k x n generator matrix after the caller's row
 * and column permutations, `bias` n symbols (NULL: zero).  1 <= k <= 15,
 * n <= 32.  Replaces CodedVecs + applyFilters (src/Candidates/LinearCodes.cpp:
 * 204-328, src/Candidates.cpp:178-188) for the LinearCode method. */
IGC_API int igc_gen_linear(igc_ctx *ctx, const uint8_t *gen, uint32_t k, uint32_t n, const uint8_t *bias,
                           const igc_filter *filter, igc_candidates *out);
/* VTCode candidates (GenerateVTCodes, src/Candidates/VTCodes.cpp): every word
 * of length n in base-4 index order with sum_{p>=1} p*[w_p >= w_{p-1}] == a
 * (mod n) and sum_p w_p == b (mod 4), then the filter. */
IGC_API int igc_gen_vt(igc_ctx *ctx, uint32_t n, int32_t a, int32_t b, const igc_filter *filter, igc_candidates *out);
/* DifferentialVTCode candidates (GenerateDifferentialVTCodes,
 * src/Candidates/DifferentialVTCodes.cpp): y_p = (w_p - w_{p+1}) mod 4,
 * y_{n-1} = w_{n-1}, kept when sum_p (p+1)*y_p == syndrome (mod 4n). */
IGC_API int igc_gen_diff_vt(igc_ctx *ctx, uint32_t n, int32_t syndrome, const igc_filter *filter,
                            igc_candidates *out);
/* AllStrings candidates: every word of length n in index order, filtered. */
IGC_API int igc_gen_all(igc_ctx *ctx, uint32_t n, const igc_filter *filter, igc_candidates *out);
/* Copies the first min(cap, count) strings of the context's latest generation
 * to host memory (cap x n bytes). */
IGC_API int igc_candidates_to_host(igc_ctx *ctx, const igc_candidates *c, uint8_t *out, uint64_t cap);

/* ---- K-means on one-hot symbol features (cluster refinement) ---- */

/* run the Lloyd iterations eagerly instead of replaying a captured hipGraph */
#define IGC_KMEANS_EAGER 1u

/* FP32 K-means of N strings of length n on their one-hot 4n-dimensional
 * features: k seeds chosen by igc_kmeans_seeds(N, k, seed), `iters` Lloyd
 * iterations (assign by squared Euclidean distance, first index on ties; update
 * = per-cluster mean, empty clusters keep their centroid), then a final
 * assignment.  assign_out[N] gets the cluster of each string; centroids_out
 * (k x 4n, may be NULL) the final centroids.  Deterministic: fixed-order sums,
 * integer histograms.  Replaces the centre assignment of the reference's
 * clustered loop (RandomCluster::fit, src/clustering/impl/random_cluster/
 * RandomCluster.cpp:42-136; src/gpu_cluster.py:16-193,287).  k*4n/2 + k <= 16384. */
IGC_API int igc_kmeans(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t k, uint32_t iters,
                       uint64_t seed, uint32_t flags, int32_t *assign_out, float *centroids_out);

/* Nearest-centroid assignment on one-hot features: out[i] = argmin_c
 * ||onehot(x_i) - C_c||^2 over k caller-provided FP32 centroids (k x 4n,
 * dimension 4p+s = symbol s at position p), ties to the lowest c; the same
 * arithmetic as igc_kmeans' assignment step.  Replaces the reference's
 * centre assignment call (src/clustering/impl/random_cluster/RandomCluster.cpp:
 * 62-116 running src/gpu_cluster.py:16-193,287). */
IGC_API int igc_assign_fp32(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, const float *centroids,
                            uint32_t k, int32_t *out);

/* The k distinct seed indices igc_kmeans starts from: a partial Fisher-Yates
 * shuffle of 0..N-1 driven by SplitMix64(seed). */
IGC_API int igc_kmeans_seeds(uint32_t N, uint32_t k, uint64_t seed, uint32_t *out);

/* ---- cluster-based refinement ---- */

#define IGC_REFINE_MAX_ROUNDS 32

typedef struct igc_refine_round {
    uint32_t candidates; /* strings entering the round */
    uint32_t clusters;   /* non-empty clusters */
    uint32_t codewords;  /* strings leaving the round */
    uint32_t pad;
    uint64_t pairs;      /* sum over clusters of N_c(N_c-1)/2 */
    float kmeans_ms;     /* device time of the K-means (HIP events) */
    float pairs_ms;      /* device time of the batched pair kernel */
    double cluster_s;    /* host wall time of the cluster step */
    double graph_s;      /* host wall time of the batched build incl. copies */
    double greedy_s;     /* host wall time of the per-cluster greedy */
} igc_refine_round;

typedef struct igc_refine_stats {
    uint32_t rounds;
    uint32_t converged;
    igc_refine_round round[IGC_REFINE_MAX_ROUNDS];
} igc_refine_stats;

/* The reference's clustered codebook loop (src/SparseMat.cpp:918-1095) on one
 * context: per round, K-means into k clusters (k raised so that N/k <= 200,000),
 * every cluster's conflict graph in one batched launch, the greedy independent
 * set per cluster on `host_threads` threads, and the cluster codebooks
 * concatenated in cluster order as the next candidates; stop when the last
 * converge_n sizes are equal or after max_rounds.  Round r seeds the K-means
 * with seed + 0x9E37*r.  out_idx (capacity N) receives the final codebook as
 * indices into syms, in order. */
IGC_API int igc_refine(igc_ctx *ctx, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist, uint32_t k,
                       uint32_t kmeans_iters, uint32_t max_rounds, uint32_t converge_n, uint64_t seed,
                       uint32_t host_threads, uint32_t *out_idx, uint32_t *out_count, igc_refine_stats *stats);

/* ---- continuous-batching server ---- */

typedef struct igc_server igc_server;

/* Result of one job: the greedy codebook (indices into the job's candidates, in
 * selection order; malloc'd, release with igc_job_free) and its timeline in
 * seconds on the server clock.  Latency = done_s - submit_s. */
typedef struct igc_job_result {
    uint64_t id;
    int status;           /* IGC_OK, or the job's error (IGC_ERR_ADMISSION when it can never fit) */
    uint32_t n_nodes;
    uint32_t codewords;
    uint32_t batch;       /* sequence number of the launch the job ran in */
    uint32_t batch_jobs;  /* jobs sharing that launch */
    uint32_t pad;
    uint32_t *order;
    uint64_t n_edges;
    double submit_s, admit_s, gpu_done_s, done_s;
    double greedy_start_s; /* a greedy worker took the job's CSR */
} igc_job_result;

/* in_use_bytes: HBM the pipelines have reserved now (their buffers plus the
 * running batches' estimates); max_in_use_bytes: the highest it has been.
 * Both stay within the budget, and the buffers within them. */
typedef struct igc_server_stats {
    uint64_t batches, jobs_done, pairs_done, queued;
    uint64_t in_use_bytes, max_in_use_bytes, max_batch_jobs;
    double gpu_busy_s;
} igc_server_info;

/* Starts a server on `device` with an HBM budget (bytes; 0 = 200 GiB) shared by
 * the launches in flight, `pipelines` GPU pipelines (1 to 4, each with its own
 * stream) and `host_threads` greedy workers.  Jobs submitted while others run
 * are admitted FIFO as soon as a pipeline is free and the budget allows, and
 * all jobs admitted together run as one batched build. */
IGC_API int igc_server_create(int device, uint64_t hbm_budget_bytes, uint32_t pipelines, uint32_t host_threads,
                              igc_server **out);
/* Queues a codebook job (the candidates are copied); returns at once.  The
 * symbols are checked here: codes above 3 (above 1 for n > 32) give
 * IGC_ERR_INVALID and queue nothing. */
IGC_API int igc_server_submit(igc_server *srv, const uint8_t *syms, uint32_t N, uint32_t n, uint32_t edit_dist,
                              uint64_t *job_id);
/* Queues a refinement job: the clustered codebook loop of igc_refine (same
 * parameters, host threads = the server's greedy threads) on N quaternary
 * strings.  It runs alone on one pipeline, which keeps the job's strings,
 * its round state and its captured round graph on the device across the
 * rounds, while the other pipeline goes on serving; the result's order holds
 * the final codebook as indices into syms, equal to igc_refine's.  No
 * counterpart in the reference (its clustered loop is a single blocking run,
 * src/SparseMat.cpp:918-1095). */
IGC_API int igc_server_submit_refine(igc_server *srv, const uint8_t *syms, uint32_t N, uint32_t n,
                                     uint32_t edit_dist, uint32_t k, uint32_t kmeans_iters, uint32_t max_rounds,
                                     uint32_t converge_n, uint64_t seed, uint64_t *job_id);
/* Waits up to timeout_ms (< 0: forever) for a job; IGC_ERR_BUSY on timeout.
 * On success the job is retired and its result handed over; when the job
 * itself failed (out->status != IGC_OK), igc_last_error() in the waiting
 * thread gives its message. */
IGC_API int igc_server_wait(igc_server *srv, uint64_t job_id, int64_t timeout_ms, igc_job_result *out);
IGC_API int igc_server_poll(igc_server *srv, uint64_t job_id, igc_job_result *out);
IGC_API int igc_job_free(igc_job_result *res);
IGC_API double igc_server_clock(igc_server *srv);
IGC_API int igc_server_stats(igc_server *srv, igc_server_info *out);
/* Drains the queue, joins the threads and releases the devices' buffers. */
IGC_API int igc_server_destroy(igc_server *srv);

/* ---- greedy independent set and outputs (host) ---- */

/* Min-degree greedy independent set of a host CSR, with the reference's exact
 * bucket semantics (AdjList::RowsBySum/MinSumRow/DelBall/DelRowCol/DecreaseSum/
 * DeleteRow, src/SparseMat.cpp:81-251; Codebook loop :698-729).  Writes the
 * selected node ids in selection order to `order` (capacity n_nodes). */
IGC_API int igc_greedy(const igc_csr *csr, uint32_t *order, uint32_t *count);
/* The same with a thread count (0: automatic, one thread below 2^24 edges, else
 * three; at most 17).  The calling thread makes every bucket move in the
 * reference's order; threads - 1 helpers find, ahead of it, each ball member's
 * neighbours still active, so the selection is the same for every count.  No
 * counterpart in the reference, whose greedy is one thread. */
IGC_API int igc_greedy_mt(const igc_csr *csr, uint32_t *order, uint32_t *count, uint32_t threads);

/* Writes the graph as the reference's edge file: little-endian int32 pairs
 * (u, v), u < v, one per edge, ascending (format of src/gpu_graph_generator.py:57-77
 * read by AdjList::FromBinaryFile, src/SparseMat.cpp:338-390). */
IGC_API int igc_write_pair_file(const igc_csr *csr, const char *path);

/* Header values of the reference's codebook file (src/Utils.cpp:311-379).
 * A clustered run (cluster_k > 0, the reference's clustering.enabled) writes
 * the clustering lines instead of the ones count, and when it converged
 * (cluster_iterations > 0) the clustering and solving times in place of the
 * fill and process times, fill_s then holding the total clustering time and
 * process_s the total solving time (src/Utils.cpp:345-368,
 * src/SparseMat.cpp:1100-1122). */
typedef struct igc_codebook_info {
    int code_len;
    int min_ed;
    int max_run;
    double min_gc;
    double max_gc;
    const char *method_line;  /* e.g. "Using Generation Method: Random (seed=1)"; may be NULL */
    int64_t candidate_num;
    int64_t matrix_ones;      /* 2*|E| */
    int threads;
    double candidates_s;
    double fill_s;
    double process_s;
    double overall_s;
    int cluster_k;              /* clustering.k; <= 0: not a clustered run */
    int cluster_iterations;     /* rounds until convergence (0: did not converge) */
    int convergence_iterations; /* clustering.convergenceIterations */
    const char *cluster_method; /* clustering.method; NULL: "" */
} igc_codebook_info;

/* Writes dir/CodeSize-%07d_CodeLen-%02d_MinED-%d.txt (src/Utils.cpp:381-418):
 * the parameter header, the '=' separator line, then one codeword per line as
 * digits '0'..'3' in selection order.  The file name is returned in name_out
 * (capacity name_cap) when name_out is not NULL. */
IGC_API int igc_write_codebook(const char *dir, const uint8_t *syms, uint32_t n, const uint32_t *order,
                               uint32_t count, const igc_codebook_info *info, char *name_out, size_t name_cap);

/* ---- diagnostics ---- */

IGC_API const char *igc_last_error(void);
/* "hip:gfx950" when a device context can be created, else "none". */
IGC_API const char *igc_backend(void);
IGC_API const char *igc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* IGC_H_ */
