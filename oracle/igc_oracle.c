/*
 * igc_oracle.c -- CPU restatement of the IndexGen conflict-graph stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The shipped library (indexgen_amd/csrc) never links or calls it.
 *
 * What it restates (reference paths relative to the IndexGen tree):
 *   orc_levenshtein      textbook Wagner-Fischer, src/Utils.cpp:15-51
 *   orc_banded           GKR/Hyyro banded bit-parallel distance with early exit,
 *                        include/EditDistance.hpp:164-279 (ComposeEqBand,
 *                        GKRHyyroSingleWordBand) and its exact fallback
 *                        MyersSingleWord :78-105
 *   orc_conflict         EditDistanceBandedAtLeast, :368-372, used as the edge
 *                        predicate in src/SparseMat.cpp:544-554
 *   orc_graph_csr        FillAdjListTH / FillAdjList, src/SparseMat.cpp:536-615:
 *                        all pairs i<j, rows striped cyclically over T threads,
 *                        merged in thread order with Set(i,j), Set(j,i)
 *   orc_greedy           AdjList min-degree greedy, src/SparseMat.cpp:81-284 and
 *                        Codebook :679-731
 *   orc_max_run, orc_gc_ok, orc_filter
 *                        MaxRun / GCContent / TestGCCont, src/Utils.cpp:420-462,
 *                        applyFilters, src/CandidateGenerator.cpp:24-50
 *
 * Symbols are bytes 0..3 (A,C,G,T), strings of equal length n.
 *
 * How each function is pinned (see tests/test_oracle.py):
 *   - orc_banded / orc_conflict: against the compiled reference's own
 *     EditDistanceBanded on golden vectors (tests/golden), and against
 *     orc_levenshtein exhaustively for n <= 6 and on random / near pairs.
 *   - orc_levenshtein: closed forms (ED(x,x)=0, ED(x,"")=|x|, single edits,
 *     Hamming upper bound, count-difference lower bound, symmetry).
 *   - orc_graph_csr + orc_greedy: against the compiled reference's
 *     FillAdjList + AdjList::ToFile + Codebook outputs (neighbour order and
 *     codebook in selection order) for several T, and the greedy invariants
 *     (independence, maximality).
 *   - filters: against brute-force definitions and the reference's double
 *     arithmetic on the boundary fractions.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Wagner-Fischer (src/Utils.cpp:15-51): one rolling row over the shorter string */
ORC_EXPORT int orc_levenshtein(const uint8_t *a, int na, const uint8_t *b, int nb)
{
    if (na > nb) {
        const uint8_t *t = a; a = b; b = t;
        int tn = na; na = nb; nb = tn;
    }
    int row[257];
    if (na > 256) return -1;
    for (int i = 0; i <= na; ++i) row[i] = i;
    for (int j = 1; j <= nb; ++j) {
        int diag = row[0];
        row[0] = j;
        for (int i = 1; i <= na; ++i) {
            int up = row[i];
            int best;
            if (a[i - 1] == b[j - 1]) {
                best = diag;
            } else {
                best = row[i - 1] < up ? row[i - 1] : up;
                if (diag < best) best = diag;
                best += 1;
            }
            row[i] = best;
            diag = up;
        }
    }
    return row[na];
}

/* ------------------------------------------------------------------ */
/* Single-word bit-parallel pieces of include/EditDistance.hpp.  Patterns are
 * limited to m <= 64 (one 64-bit block), which covers every length this
 * stage accepts. */

/* Match masks: bit i of peq[c] set iff pattern[i] == c (PeqIndex, :10-37). */
static void peq_build(const uint8_t *pat, int m, uint64_t peq[256])
{
    memset(peq, 0, 256 * sizeof(uint64_t));
    for (int i = 0; i < m; ++i) peq[pat[i]] |= 1ULL << i;
}

/* Exact Myers on one word (MyersSingleWord, :78-105). */
static int myers_word(const uint64_t peq[256], int m, const uint8_t *text, int n)
{
    if (m == 0) return n;
    uint64_t pv = ~0ULL, mv = 0ULL;
    int score = m;
    int top = (m - 1) & 63;
    for (int j = 0; j < n; ++j) {
        uint64_t eq = peq[text[j]];
        uint64_t x = eq | mv;
        uint64_t d0 = (((x & pv) + pv) ^ pv) | x;
        uint64_t hn = pv & d0;
        uint64_t hp = mv | ~(pv | d0);
        uint64_t x2 = (hp << 1) | 1ULL;
        mv = x2 & d0;
        pv = (hn << 1) | ~(x2 | d0);
        score += (int)((hp >> top) & 1ULL) - (int)((hn >> top) & 1ULL);
    }
    return score;
}

/* Eq word of the diagonal band whose most significant bit is pattern row
 * j + c (ComposeEqBand, :164-190), specialised to one pattern block. */
static uint64_t band_eq(const uint64_t peq[256], uint8_t ch, int j, int c)
{
    const int itop = j + c;
    const int blk = itop >> 6;
    const int off = itop & 63;
    const uint64_t w = peq[ch];
    uint64_t eq = 0ULL;
    if (blk == 0) {
        uint64_t keep = (off == 63) ? ~0ULL : ((1ULL << (off + 1)) - 1ULL);
        eq |= (w & keep) << (63 - off);
    }
    if (off < 63 && blk - 1 == 0) {
        const int need = 63 - off;
        eq |= (w & (~0ULL << (64 - need))) >> (off + 1);
    }
    return eq;
}

/* Thresholded distance: the exact distance if it is <= k, else k + 1
 * (GKRHyyroSingleWordBand, :194-279).  text has length n, pattern length m. */
ORC_EXPORT int orc_banded(const uint8_t *text, int n, const uint8_t *pat, int m, int k)
{
    if (m == 0) return n <= k ? n : k + 1;
    if (n == 0) return m <= k ? m : k + 1;
    if (m > 64) return -2; /* outside the single-block restatement */
    int diff = n - m;
    if (diff < 0) diff = -diff;
    if (k < diff) return k + 1;

    uint64_t peq[256];
    peq_build(pat, m, peq);

    const int shortest = m < n ? m : n;
    if (k >= shortest) {
        int d = myers_word(peq, m, text, n);
        return d <= k ? d : k + 1;
    }

    const int c = (k - n + m) / 2;
    int lv = c + (k + n - m) / 2 + 1;
    if (lv > m) lv = m;
    if (lv > 64 || lv == m) {
        int d = myers_word(peq, m, text, n);
        return d <= k ? d : k + 1;
    }

    /* top c+1 bits set */
    uint64_t vp = (~0ULL >> (64 - (c + 1))) << (64 - (c + 1));
    uint64_t vn = 0ULL;
    int score = c;
    const int late_from = m - c;
    int s = 62;
    const int late_cols = (n - late_from) > 0 ? (n - late_from) : 0;

    for (int j = 0; j < n; ++j) {
        uint64_t eq = band_eq(peq, text[j], j, c);
        uint64_t x = eq | vn;
        uint64_t d0 = (((x & vp) + vp) ^ vp) | x;
        uint64_t hn = vp & d0;
        uint64_t hp = vn | ~(d0 | vp);
        uint64_t sh = d0 >> 1;
        vn = sh & hp;
        vp = hn | ~(sh | hp);
        if (j < late_from) {
            score += 1 - (int)((d0 >> 63) & 1ULL);
        } else {
            score += (int)((hp >> s) & 1ULL);
            score -= (int)((hn >> s) & 1ULL);
            --s;
        }
        int rem = (j + 1 < late_from) ? late_cols : (n - (j + 1));
        if (score - rem > k) return k + 1;
    }
    return score <= k ? score : k + 1;
}

/* Edge predicate of the all-pairs loop: pattern = row string a, text = column
 * string b, edge iff the banded distance capped at minED-1 is below minED
 * (EditDistanceBandedAtLeast, :368-372; src/SparseMat.cpp:544-554). */
ORC_EXPORT int orc_conflict(const uint8_t *a, const uint8_t *b, int n, int min_ed)
{
    int d = orc_banded(b, n, a, n, min_ed - 1);
    return d < min_ed;
}

/* ------------------------------------------------------------------ */
/* All-pairs conflict graph with the reference's neighbour order. */

typedef struct {
    const uint8_t *syms;
    int N, n, min_ed;
    int worker, workers;
    uint32_t **row_js;   /* per row i: ascending j>i with an edge */
    uint32_t *row_cnt;
} scan_job;

static void *scan_rows(void *arg)
{
    scan_job *jb = (scan_job *)arg;
    const int N = jb->N, n = jb->n;
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(N > 0 ? N : 1));
    for (int i = jb->worker; i < N; i += jb->workers) {
        const uint8_t *a = jb->syms + (size_t)i * n;
        uint32_t cnt = 0;
        for (int j = i + 1; j < N; ++j) {
            if (orc_conflict(a, jb->syms + (size_t)j * n, n, jb->min_ed)) buf[cnt++] = (uint32_t)j;
        }
        jb->row_cnt[i] = cnt;
        jb->row_js[i] = NULL;
        if (cnt) {
            jb->row_js[i] = (uint32_t *)malloc(sizeof(uint32_t) * cnt);
            memcpy(jb->row_js[i], buf, sizeof(uint32_t) * cnt);
        }
    }
    free(buf);
    return NULL;
}

typedef struct {
    int N;
    uint32_t **row_js;
    uint32_t *row_cnt;
} orc_pairs;

static int compute_pairs(const uint8_t *syms, int N, int n, int min_ed, int workers, orc_pairs *out)
{
    out->N = N;
    out->row_js = (uint32_t **)calloc((size_t)(N > 0 ? N : 1), sizeof(uint32_t *));
    out->row_cnt = (uint32_t *)calloc((size_t)(N > 0 ? N : 1), sizeof(uint32_t));
    if (workers < 1) workers = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * workers);
    scan_job *jobs = (scan_job *)malloc(sizeof(scan_job) * workers);
    for (int w = 0; w < workers; ++w) {
        jobs[w] = (scan_job){syms, N, n, min_ed, w, workers, out->row_js, out->row_cnt};
        pthread_create(&th[w], NULL, scan_rows, &jobs[w]);
    }
    for (int w = 0; w < workers; ++w) pthread_join(th[w], NULL);
    free(th);
    free(jobs);
    return 0;
}

static void free_pairs(orc_pairs *p)
{
    for (int i = 0; i < p->N; ++i) free(p->row_js[i]);
    free(p->row_js);
    free(p->row_cnt);
}

/* Number of undirected edges (i<j) of the conflict graph. */
ORC_EXPORT int64_t orc_edge_count(const uint8_t *syms, int N, int n, int min_ed, int workers)
{
    orc_pairs p;
    compute_pairs(syms, N, n, min_ed, workers, &p);
    int64_t e = 0;
    for (int i = 0; i < N; ++i) e += p.row_cnt[i];
    free_pairs(&p);
    return e;
}

/* Symmetric adjacency in the order the reference's T-thread merge produces
 * (src/SparseMat.cpp:599-609): thread t owns rows t, t+T, ...; thread lists are
 * walked in thread order and each pair (i,j) appends j to row i and i to row j.
 * T = 1 gives ascending neighbour lists.
 *
 * Call with col == NULL to get only row_ptr (length N+1).  Returns |E|. */
ORC_EXPORT int64_t orc_graph_csr(const uint8_t *syms, int N, int n, int min_ed, int T, int workers,
                                 uint64_t *row_ptr, uint32_t *col)
{
    if (T < 1) T = 1;
    orc_pairs p;
    compute_pairs(syms, N, n, min_ed, workers, &p);
    uint64_t *deg = (uint64_t *)calloc((size_t)N + 1, sizeof(uint64_t));
    int64_t e = 0;
    for (int i = 0; i < N; ++i) {
        e += p.row_cnt[i];
        deg[i] += p.row_cnt[i];
        for (uint32_t t = 0; t < p.row_cnt[i]; ++t) deg[p.row_js[i][t]] += 1;
    }
    row_ptr[0] = 0;
    for (int i = 0; i < N; ++i) row_ptr[i + 1] = row_ptr[i] + deg[i];
    if (col) {
        uint64_t *fill = deg; /* reuse as per-row cursor */
        for (int i = 0; i < N; ++i) fill[i] = row_ptr[i];
        for (int t = 0; t < T; ++t) {
            for (int i = t; i < N; i += T) {
                for (uint32_t q = 0; q < p.row_cnt[i]; ++q) {
                    uint32_t j = p.row_js[i][q];
                    col[fill[i]++] = j;
                    col[fill[j]++] = (uint32_t)i;
                }
            }
        }
    }
    free(deg);
    free_pairs(&p);
    return e;
}

/* ------------------------------------------------------------------ */
/* Min-degree greedy independent set (AdjList, src/SparseMat.cpp:81-284;
 * Codebook, :679-731).  Buckets hold nodes by current degree; the pick is the
 * last node of the lowest non-empty bucket; removals swap the bucket's last
 * node into the hole; a node whose degree drops is appended to bucket d-1.
 * Writes the selected nodes in selection order and returns their count. */
ORC_EXPORT int64_t orc_greedy(int N, const uint64_t *row_ptr, const uint32_t *col, int32_t *order)
{
    if (N <= 0) return 0;
    int32_t *degree = (int32_t *)malloc(sizeof(int32_t) * N);
    int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * N);
    uint8_t *deleted = (uint8_t *)calloc(N, 1);
    int maxd = 0;
    for (int i = 0; i < N; ++i) {
        degree[i] = (int32_t)(row_ptr[i + 1] - row_ptr[i]);
        if (degree[i] > maxd) maxd = degree[i];
    }
    /* One growable array per degree bucket. */
    int nb = maxd + 1;
    int32_t **bk = (int32_t **)calloc(nb, sizeof(int32_t *));
    int32_t *blen = (int32_t *)calloc(nb, sizeof(int32_t));
    int32_t *bcap = (int32_t *)calloc(nb, sizeof(int32_t));
    #define BK_PUSH(d, v) do { \
        if (blen[d] == bcap[d]) { bcap[d] = bcap[d] ? 2 * bcap[d] : 8; \
            bk[d] = (int32_t *)realloc(bk[d], sizeof(int32_t) * bcap[d]); } \
        pos[v] = blen[d]; bk[d][blen[d]++] = (v); } while (0)

    int tracker = nb;
    for (int i = 0; i < N; ++i) {
        int d = degree[i];
        BK_PUSH(d, i);
        if (d < tracker) tracker = d;
    }
    int active = N;
    int64_t picked = 0;

    /* DeleteRow: swap-remove from bucket d, advance the tracker past empties. */
    #define BK_REMOVE(d, v) do { \
        int32_t p_ = pos[v]; int32_t last_ = bk[d][blen[d] - 1]; \
        bk[d][p_] = last_; pos[last_] = p_; blen[d]--; \
        if ((d) == tracker && blen[d] == 0) { \
            while (tracker < nb && blen[tracker] == 0) tracker++; } } while (0)

    int32_t *todel = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxd + 1));
    while (active > 0) {
        int v = bk[tracker][blen[tracker] - 1];
        order[picked++] = v;
        /* DelBall: active neighbours in list order, then v itself. */
        int nt = 0;
        for (uint64_t q = row_ptr[v]; q < row_ptr[v + 1]; ++q) {
            uint32_t u = col[q];
            if (!deleted[u]) todel[nt++] = (int32_t)u;
        }
        todel[nt++] = v;
        for (int t = 0; t < nt; ++t) {
            int x = todel[t];
            if (deleted[x]) continue;
            /* DelRowCol: lower every active neighbour's degree in list order. */
            for (uint64_t q = row_ptr[x]; q < row_ptr[x + 1]; ++q) {
                uint32_t u = col[q];
                if (deleted[u]) continue;
                int du = degree[u];
                if (du > 0) {
                    BK_REMOVE(du, (int)u);
                    degree[u] = du - 1;
                    BK_PUSH(du - 1, (int)u);
                    if (du - 1 < tracker) tracker = du - 1;
                }
            }
            int dx = degree[x];
            BK_REMOVE(dx, x);
            deleted[x] = 1;
            active--;
        }
    }
    #undef BK_PUSH
    #undef BK_REMOVE
    for (int d = 0; d < nb; ++d) free(bk[d]);
    free(bk); free(blen); free(bcap); free(todel);
    free(degree); free(pos); free(deleted);
    return picked;
}

/* ------------------------------------------------------------------ */
/* Filters (src/Utils.cpp:420-462; src/CandidateGenerator.cpp:24-50). */
ORC_EXPORT int orc_max_run(const uint8_t *s, int n)
{
    if (n <= 0) return 0;
    int best = 1, run = 1;
    for (int i = 1; i < n; ++i) {
        if (s[i] == s[i - 1]) {
            run++;
        } else {
            run = 1;
        }
        if (run > best) best = run;
    }
    return best;
}

ORC_EXPORT int orc_gc_ok(const uint8_t *s, int n, double min_gc, double max_gc)
{
    int gc = 0;
    for (int i = 0; i < n; ++i) gc += (s[i] == 1 || s[i] == 2);
    double frac = (double)gc / (double)n;
    return frac >= min_gc && frac <= max_gc;
}

/* Keeps strings passing both filters, in order; returns the kept count and
 * writes kept indices to keep[].  max_run <= 0 disables the run filter; the GC
 * filter is active when either bound is positive (applyFilters). */
ORC_EXPORT int64_t orc_filter(const uint8_t *syms, int64_t N, int n, int max_run, double min_gc,
                              double max_gc, int64_t *keep)
{
    const int use_run = max_run > 0;
    const int use_gc = (min_gc > 0.0 || max_gc > 0.0);
    int64_t k = 0;
    for (int64_t i = 0; i < N; ++i) {
        const uint8_t *s = syms + (size_t)i * n;
        int ok = 1;
        if (use_run && orc_max_run(s, n) > max_run) ok = 0;
        if (ok && use_gc && !orc_gc_ok(s, n, min_gc, max_gc)) ok = 0;
        if (ok) keep[k++] = i;
    }
    return k;
}

/* ------------------------------------------------------------------ */
/* Bulk checks used by the tests. */

/* Over all ordered pairs of length-n strings (n <= 8) and every k in
 * [0, kmax]: counts cases where orc_banded(text, pattern, k) differs from
 * min(orc_levenshtein, k+1).  Exhaustive pin of the banded restatement. */
ORC_EXPORT int64_t orc_check_exhaustive(int n, int kmax)
{
    if (n < 1 || n > 8) return -1;
    const int64_t S = 1LL << (2 * n);
    uint8_t a[8], b[8];
    int64_t bad = 0;
    for (int64_t x = 0; x < S; ++x) {
        for (int p = 0; p < n; ++p) a[p] = (uint8_t)((x >> (2 * p)) & 3);
        for (int64_t y = 0; y < S; ++y) {
            for (int p = 0; p < n; ++p) b[p] = (uint8_t)((y >> (2 * p)) & 3);
            int wf = orc_levenshtein(a, n, b, n);
            for (int k = 0; k <= kmax; ++k) {
                int want = wf <= k ? wf : k + 1;
                if (orc_banded(b, n, a, n, k) != want) bad++;
            }
        }
    }
    return bad;
}

/* Neighbours of row i (all j != i with an edge), ascending; returns count. */
ORC_EXPORT int64_t orc_row_neighbors(const uint8_t *syms, int N, int n, int min_ed, int i, uint32_t *out)
{
    int64_t c = 0;
    const uint8_t *a = syms + (size_t)i * n;
    for (int j = 0; j < N; ++j) {
        if (j == i) continue;
        const uint8_t *b = syms + (size_t)j * n;
        /* the reference tests row min(i,j) as pattern against max(i,j) as text */
        int e = (i < j) ? orc_conflict(a, b, n, min_ed) : orc_conflict(b, a, n, min_ed);
        if (e) out[c++] = (uint32_t)j;
    }
    return c;
}

/* out[i*N + j] = edge predicate for rows i (pattern) and j (text), i != j; the
 * diagonal is 0.  For exhaustive comparisons with the GPU predicate. */
ORC_EXPORT void orc_conflict_matrix(const uint8_t *syms, int N, int n, int min_ed, uint8_t *out)
{
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            out[(size_t)i * N + j] = (uint8_t)(i != j && orc_conflict(syms + (size_t)i * n, syms + (size_t)j * n, n, min_ed));
}

/* out[r] = edge predicate for the pair (A[r] pattern, B[r] text). */
ORC_EXPORT void orc_conflict_pairs(const uint8_t *A, const uint8_t *B, int M, int n, int min_ed, uint8_t *out)
{
    for (int r = 0; r < M; ++r) out[r] = (uint8_t)orc_conflict(A + (size_t)r * n, B + (size_t)r * n, n, min_ed);
}

/* ------------------------------------------------------------------ */
/* FP32 K-means on one-hot symbol features: the CPU statement of the GPU
 * kernels' arithmetic (no reference counterpart: the reference assigns to
 * random centres by edit distance, src/clustering/impl/random_cluster/
 * RandomCluster.cpp:42-136).  Seeds: partial Fisher-Yates over 0..N-1 driven
 * by SplitMix64(seed).  Every float sum runs in the same order as on the GPU;
 * this file is compiled with -ffp-contract=off. */
static uint64_t splitmix_next(uint64_t *st)
{
    *st += 0x9E3779B97F4A7C15ULL;
    uint64_t z = *st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

ORC_EXPORT void orc_kmeans_seeds(int N, int k, uint64_t seed, uint32_t *out)
{
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(N > 0 ? N : 1));
    for (int i = 0; i < N; ++i) idx[i] = (uint32_t)i;
    uint64_t st = seed;
    for (int i = 0; i < k; ++i) {
        uint64_t z = splitmix_next(&st);
        uint32_t j = (uint32_t)i + (uint32_t)(z % (uint64_t)(N - i));
        uint32_t t = idx[i]; idx[i] = idx[j]; idx[j] = t;
        out[i] = idx[i];
    }
    free(idx);
}

static void km_assign(const uint8_t *syms, int N, int n, int k, const float *C, const float *cn2, int32_t *assign)
{
    const int D = 4 * n;
    for (int i = 0; i < N; ++i) {
        const uint8_t *s = syms + (size_t)i * n;
        float best = 0.0f;
        int bi = -1;
        for (int c = 0; c < k; ++c) {
            const float *row = C + (size_t)c * D;
            float acc = 0.0f;
            for (int p = 0; p < n; ++p) acc = acc + row[4 * p + s[p]];
            float twice = 2.0f * acc;
            float dist = cn2[c] - twice;
            if (bi < 0 || dist < best) { best = dist; bi = c; }
        }
        assign[i] = bi;
    }
}

/* Nearest-centroid assignment with caller-provided centroids: cn2 in dimension
 * order, then the same argmin as the Lloyd iterations above. */
ORC_EXPORT void orc_assign_fp32(const uint8_t *syms, int N, int n, const float *C, int k, int32_t *assign)
{
    const int D = 4 * n;
    float *cn2 = (float *)calloc((size_t)k, sizeof(float));
    for (int c = 0; c < k; ++c) {
        float acc = 0.0f;
        for (int d = 0; d < D; ++d) {
            float v = C[(size_t)c * D + d];
            float sq = v * v;
            acc = acc + sq;
        }
        cn2[c] = acc;
    }
    km_assign(syms, N, n, k, C, cn2, assign);
    free(cn2);
}

ORC_EXPORT void orc_kmeans(const uint8_t *syms, int N, int n, int k, int iters, uint64_t seed, int32_t *assign,
                           float *C_out)
{
    const int D = 4 * n;
    float *C = (float *)calloc((size_t)k * D, sizeof(float));
    float *cn2 = (float *)calloc((size_t)k, sizeof(float));
    uint32_t *cnt = (uint32_t *)calloc((size_t)k * D, sizeof(uint32_t));
    uint32_t *size = (uint32_t *)calloc((size_t)k, sizeof(uint32_t));
    uint32_t *seeds = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)k);
    orc_kmeans_seeds(N, k, seed, seeds);
    for (int c = 0; c < k; ++c) {
        const uint8_t *s = syms + (size_t)seeds[c] * n;
        for (int p = 0; p < n; ++p) C[(size_t)c * D + 4 * p + s[p]] = 1.0f;
        cn2[c] = (float)n;
    }
    for (int it = 0; it < iters; ++it) {
        km_assign(syms, N, n, k, C, cn2, assign);
        memset(cnt, 0, sizeof(uint32_t) * (size_t)k * D);
        memset(size, 0, sizeof(uint32_t) * (size_t)k);
        for (int i = 0; i < N; ++i) {
            const uint8_t *s = syms + (size_t)i * n;
            for (int p = 0; p < n; ++p) cnt[(size_t)assign[i] * D + 4 * p + s[p]]++;
            size[assign[i]]++;
        }
        for (int c = 0; c < k; ++c) {
            if (size[c])
                for (int d = 0; d < D; ++d) C[(size_t)c * D + d] = (float)cnt[(size_t)c * D + d] / (float)size[c];
            float acc = 0.0f;
            for (int d = 0; d < D; ++d) {
                float v = C[(size_t)c * D + d];
                float sq = v * v;
                acc = acc + sq;
            }
            cn2[c] = acc;
        }
    }
    km_assign(syms, N, n, k, C, cn2, assign);
    if (C_out) memcpy(C_out, C, sizeof(float) * (size_t)k * D);
    free(C); free(cn2); free(cnt); free(size); free(seeds);
}

/* ------------------------------------------------------------------ */
/* Candidate generators with applyFilters (src/CandidateGenerator.cpp:24-50).
 * Each enumerates its words in the reference's order, keeps those that pass
 * the generator's test and the filters (orc_filter's rules), writes the first
 * `cap` kept strings to out (N x n symbols) and returns how many were kept.
 *
 *   orc_gen_linear  CodedVecs (src/Candidates/LinearCodes.cpp:204-328): every
 *                   message m of length k in NextBase4 order (src/Utils.cpp:
 *                   525-545; last digit fastest), codeword m*G + bias over
 *                   GF(4) (src/Candidates/GF4.cpp:15-60: addition is XOR of
 *                   the labels, multiplication the table below); G is the
 *                   generator matrix after the caller's permutations.
 *   orc_gen_vt      GenerateVTCodes (src/Candidates/VTCodes.cpp): words in
 *                   base-4 index order (last symbol least significant), kept
 *                   when sum over p >= 1 of p*[w[p] >= w[p-1]] == a (mod n) and
 *                   sum of symbols == b (mod 4).
 *   orc_gen_dvt     GenerateDifferentialVTCodes (src/Candidates/
 *                   DifferentialVTCodes.cpp): y[p] = (w[p] - w[p+1]) mod 4 for
 *                   p < n-1, y[n-1] = w[n-1]; kept when sum (p+1)*y[p] == s
 *                   (mod 4n).
 *   orc_gen_all     AllStringsGenerator: every word in index order.
 * Pinned by tests/golden/gen_golden.npz, written by the compiled reference's
 * own generators (tests/golden/make_golden.py). */
static const uint8_t orc_gf4_mul[4][4] = {{0, 0, 0, 0}, {0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}};

static int orc_keep(const uint8_t *w, int n, int max_run, double min_gc, double max_gc)
{
    const int use_run = max_run > 0;
    const int use_gc = (min_gc > 0.0 || max_gc > 0.0);
    if (use_run && orc_max_run(w, n) > max_run) return 0;
    if (use_gc && !orc_gc_ok(w, n, min_gc, max_gc)) return 0;
    return 1;
}

static void orc_index_word(uint64_t t, int n, uint8_t *w)
{
    for (int p = n - 1; p >= 0; --p) {
        w[p] = (uint8_t)(t & 3u);
        t >>= 2;
    }
}

ORC_EXPORT int64_t orc_gen_linear(const uint8_t *G, int k, int n, const uint8_t *bias, int max_run, double min_gc,
                                  double max_gc, uint8_t *out, int64_t cap)
{
    if (k < 1 || k > 15 || n < 1 || n > 64) return -1;
    uint8_t m[16], w[64];
    int64_t kept = 0;
    const uint64_t total = 1ull << (2 * k);
    for (uint64_t t = 0; t < total; ++t) {
        orc_index_word(t, k, m);
        for (int j = 0; j < n; ++j) {
            uint8_t c = bias ? bias[j] : 0;
            for (int i = 0; i < k; ++i) c ^= orc_gf4_mul[m[i]][G[(size_t)i * n + j]];
            w[j] = c;
        }
        if (!orc_keep(w, n, max_run, min_gc, max_gc)) continue;
        if (kept < cap) memcpy(out + (size_t)kept * n, w, (size_t)n);
        ++kept;
    }
    return kept;
}

ORC_EXPORT int64_t orc_gen_vt(int n, int a, int b, int max_run, double min_gc, double max_gc, uint8_t *out,
                              int64_t cap)
{
    if (n < 1 || n > 15) return -1;
    uint8_t w[16];
    int64_t kept = 0;
    const long long am = ((a % n) + n) % n, bm = ((b % 4) + 4) % 4;
    const uint64_t total = 1ull << (2 * n);
    for (uint64_t t = 0; t < total; ++t) {
        orc_index_word(t, n, w);
        long long alpha = 0, sum = 0;
        for (int p = 1; p < n; ++p)
            if (w[p] >= w[p - 1]) alpha += p;
        for (int p = 0; p < n; ++p) sum += w[p];
        if (alpha % n != am || sum % 4 != bm) continue;
        if (!orc_keep(w, n, max_run, min_gc, max_gc)) continue;
        if (kept < cap) memcpy(out + (size_t)kept * n, w, (size_t)n);
        ++kept;
    }
    return kept;
}

ORC_EXPORT int64_t orc_gen_dvt(int n, int s, int max_run, double min_gc, double max_gc, uint8_t *out, int64_t cap)
{
    if (n < 1 || n > 15) return -1;
    uint8_t w[16];
    int64_t kept = 0;
    const long long mod = 4ll * n, sm = ((s % mod) + mod) % mod;
    const uint64_t total = 1ull << (2 * n);
    for (uint64_t t = 0; t < total; ++t) {
        orc_index_word(t, n, w);
        long long syn = 0;
        for (int p = 0; p < n; ++p) {
            const int y = p < n - 1 ? (((int)w[p] - (int)w[p + 1]) % 4 + 4) % 4 : w[n - 1];
            syn += (long long)(p + 1) * y;
        }
        if (syn % mod != sm) continue;
        if (!orc_keep(w, n, max_run, min_gc, max_gc)) continue;
        if (kept < cap) memcpy(out + (size_t)kept * n, w, (size_t)n);
        ++kept;
    }
    return kept;
}

ORC_EXPORT int64_t orc_gen_all(int n, int max_run, double min_gc, double max_gc, uint8_t *out, int64_t cap)
{
    if (n < 1 || n > 15) return -1;
    uint8_t w[16];
    int64_t kept = 0;
    const uint64_t total = 1ull << (2 * n);
    for (uint64_t t = 0; t < total; ++t) {
        orc_index_word(t, n, w);
        if (!orc_keep(w, n, max_run, min_gc, max_gc)) continue;
        if (kept < cap) memcpy(out + (size_t)kept * n, w, (size_t)n);
        ++kept;
    }
    return kept;
}
