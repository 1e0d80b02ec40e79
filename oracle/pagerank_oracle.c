/*
 * pagerank_oracle.c -- CPU restatement of NetworKit's PageRank (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the oracle: tests, the package's smoke check and the benchmark's
 * `cpu_baseline` leg load it (through ctypes) to check or time against.  The product path
 * never links, loads or calls it; the GPU library fails loudly when its own HIP build is
 * missing instead of falling back to this code.
 *
 * What it restates (all citations are paths in kit-parco/networkit v9.0):
 *   - PageRank::run                      networkit/cpp/centrality/PageRank.cpp:20-104
 *       init / warm start                 :22-27   (scoreData.resize(z, 1/n) keeps old values)
 *       weighted out-degree               :29-30   -> Graph::weightedDegree, Graph.cpp:427-455,539
 *       pull sweep over in-edges          :56-66   (pr[u] += x[v]*w/deg[v]; pr[u] = pr[u]*damp + tele)
 *       convergence test                  :34-51,68-69 (checks iterations >= maxIterations first)
 *       post-normalisation                :75-92
 *       maximum()                         :95-102 (seeded with scoreData[0], then atomic_max)
 *   - the node loops                      include/networkit/graph/Graph.hpp:1769-1809 (skip !exists,
 *                                          guided schedule), reductions :2123-2134
 *   - in-edge semantics                   Graph.hpp:1918-1938 (directed: inEdges[u]; undirected:
 *                                          outEdges[u], self-loop stored once, Graph.cpp:577-580)
 *   - ranking order                       networkit/cpp/centrality/Centrality.cpp:29-41
 *
 * Layout.  NetworKit keeps per-node std::vector<node> (uint64 ids) and std::vector<double>
 * weights.  The oracle takes the same element types flattened into offsets:
 * in_ptr[z+1] (uint64), in_src[m] (uint64), in_w[m] (double, NULL when unweighted), exists[z]
 * (uint8, NULL when every slot exists).  For an undirected graph the "in-lists" are the
 * adjacency lists (each edge appears in both endpoint lists, a self-loop once).
 *
 * Sink handling.  The reference has no option: mass sent to a node without out-edges is lost
 * every sweep and only the final normalisation rescales it (sink_mode 0).  sink_mode 1
 * ("distribute sinks", not in the reference) adds damp * S / n to every node, where
 * S = sum of scores of existing nodes with out-degree 0 before the sweep; this is the
 * Google-matrix fixed point and is pinned by closed forms in the tests, not by the reference.
 *
 * Parity pinning (see the tests that load this file):
 *   oracle_pagerank       pinned by the reference's known answers (8-node graph +-1e-4,
 *                         CentralityGTest.cpp:582-664; wiki-Vote top-1 in 4 variants,
 *                         :484-539), by a sparse direct solve of (I - dP)x = (1-d)/n 1, by
 *                         dense power iteration on tiny graphs, and by invariants (uniform
 *                         scores on regular graphs, equal scores on automorphic nodes).
 *   oracle_ranking        pinned by a brute-force sort in the tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_L1 0
#define ORACLE_L2 1

/* Weighted out-degree per slot: sum of the weights of the entries whose source is v.  For a
 * directed graph these are the out-edges of v; for an undirected graph the adjacency of v.
 * Mirrors Graph::computeWeightedDegree(u, inDegree=false, countSelfLoopsTwice=false),
 * Graph.cpp:427-455.  The unweighted out-count (Graph::degree, Graph.hpp:1207) goes to cnt. */
void oracle_out_degrees(uint64_t z, const uint64_t *in_ptr, const uint64_t *in_src,
                        const double *in_w, double *deg, uint64_t *cnt) {
    int threads = 1;
#ifdef _OPENMP
    threads = omp_get_max_threads();
#endif
    /* The reference sums each node's out-list on its own thread.  The oracle holds in-lists,
     * so in parallel each thread adds the entries of a fixed range of rows (static schedule)
     * into a private array, and the private arrays are added in thread order: the result
     * depends on the thread count only.  Counts are exact either way.  Large graphs whose
     * private arrays would not fit a 16 GB budget are summed on one thread. */
    double *part = NULL;
    if (threads > 1 && (double)z * (double)threads * 8.0 <= 16e9)
        part = (double *)calloc((size_t)z * (size_t)threads, sizeof(double));
    for (uint64_t v = 0; v < z; ++v) {
        deg[v] = 0.0;
        if (cnt) cnt[v] = 0;
    }
    if (!part) {
        for (uint64_t u = 0; u < z; ++u) {
            for (uint64_t e = in_ptr[u]; e < in_ptr[u + 1]; ++e) {
                uint64_t v = in_src[e];
                deg[v] += in_w ? in_w[e] : 1.0;
                if (cnt) cnt[v] += 1;
            }
        }
        return;
    }
#pragma omp parallel num_threads(threads)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double *mine = part + (size_t)t * z;
#pragma omp for schedule(static)
        for (int64_t u = 0; u < (int64_t)z; ++u) {
            for (uint64_t e = in_ptr[u]; e < in_ptr[u + 1]; ++e) {
                const uint64_t v = in_src[e];
                mine[v] += in_w ? in_w[e] : 1.0;
                if (cnt) __atomic_fetch_add(&cnt[v], 1, __ATOMIC_RELAXED);
            }
        }
#pragma omp for schedule(static)
        for (int64_t v = 0; v < (int64_t)z; ++v) {
            double s = 0.0;
            for (int k = 0; k < threads; ++k) s += part[(size_t)k * z + v];
            deg[v] = s;
        }
    }
    free(part);
}

/* PageRank::run.  scores[z] holds the start vector on entry (the caller applies the
 * reference's warm-start rule: a fresh object starts at 1/n, a re-run keeps the previous
 * scores) and the final scores on exit.  Returns 0 on success, -1 on bad input.
 *
 * teleport (not in the reference, personalised PageRank): NULL gives the reference's
 * (1 - damp) / n.  Otherwise teleport[z] is a distribution t over the slots and node u
 * receives (1 - damp) * t[u] instead; distributed sink mass stays uniform (d S / n).  With t
 * uniform over the existing nodes this is the reference loop again. */
int oracle_pagerank_tv(uint64_t z, const uint8_t *exists, const uint64_t *in_ptr,
                       const uint64_t *in_src, const double *in_w, double damp, double tol,
                       int norm, uint64_t max_iterations, int normalized, int sink_mode,
                       int num_threads, const double *teleport, double *scores,
                       uint64_t *iterations_out, double *maximum_out) {
    if (!in_ptr || !scores) return -1;
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#else
    (void)num_threads;
#endif
    uint64_t n = 0;
    for (uint64_t u = 0; u < z; ++u) n += (!exists || exists[u]) ? 1 : 0;
    if (n == 0) {
        if (iterations_out) *iterations_out = 0;
        if (maximum_out) *maximum_out = normalized ? scores[0] : 1.0;
        return 0;
    }
    const double tele_uniform = (1.0 - damp) / (double)n;

    double *deg = (double *)malloc(sizeof(double) * (z ? z : 1));
    uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) * (z ? z : 1));
    double *pr = (double *)malloc(sizeof(double) * (z ? z : 1));
    if (!deg || !cnt || !pr) {
        free(deg); free(cnt); free(pr);
        return -1;
    }
    oracle_out_degrees(z, in_ptr, in_src, in_w, deg, cnt);
    memcpy(pr, scores, sizeof(double) * z);

    double *x = scores; /* "scoreData" */
    double *y = pr;     /* "pr" */
    uint64_t it = 0;
    int converged = 0;
    do {
        double sink_term = 0.0;
        if (sink_mode == 1) {
            double s = 0.0;
#pragma omp parallel for reduction(+ : s)
            for (int64_t u = 0; u < (int64_t)z; ++u)
                if ((!exists || exists[u]) && cnt[u] == 0) s += x[u];
            sink_term = damp * s / (double)n;
        }
#pragma omp parallel for schedule(guided)
        for (int64_t u = 0; u < (int64_t)z; ++u) {
            if (exists && !exists[u]) continue;
            double acc = 0.0;
            for (uint64_t e = in_ptr[u]; e < in_ptr[u + 1]; ++e) {
                uint64_t v = in_src[e];
                double w = in_w ? in_w[e] : 1.0;
                acc += x[v] * w / deg[v];
            }
            acc *= damp;
            acc += teleport ? (1.0 - damp) * teleport[u] : tele_uniform;
            acc += sink_term;
            y[u] = acc;
        }
        ++it;
        if (it >= max_iterations) {
            converged = 1;
        } else {
            double s = 0.0;
            if (norm == ORACLE_L2) {
#pragma omp parallel for reduction(+ : s)
                for (int64_t u = 0; u < (int64_t)z; ++u) {
                    if (exists && !exists[u]) continue;
                    double d = x[u] - y[u];
                    s += d * d;
                }
                converged = sqrt(s) <= tol;
            } else {
#pragma omp parallel for reduction(+ : s)
                for (int64_t u = 0; u < (int64_t)z; ++u) {
                    if (exists && !exists[u]) continue;
                    s += fabs(x[u] - y[u]);
                }
                converged = s <= tol;
            }
        }
        double *t = x; x = y; y = t; /* std::swap(pr, scoreData) */
    } while (!converged);

    /* after an odd number of swaps the newest vector lives in pr: copy it back */
    if (x != scores) memcpy(scores, x, sizeof(double) * z);

    if (!normalized) {
        double s = 0.0;
#pragma omp parallel for reduction(+ : s)
        for (int64_t u = 0; u < (int64_t)z; ++u)
            if (!exists || exists[u]) s += scores[u];
#pragma omp parallel for
        for (int64_t u = 0; u < (int64_t)z; ++u)
            if (!exists || exists[u]) scores[u] /= s;
    } else {
        double s = 0.0;
#pragma omp parallel for reduction(+ : s)
        for (int64_t u = 0; u < (int64_t)z; ++u)
            if ((!exists || exists[u]) && cnt[u] == 0) s += scores[u];
        const double norm_factor = (1.0 / (double)n) * ((1.0 - damp) + damp * s);
#pragma omp parallel for
        for (int64_t u = 0; u < (int64_t)z; ++u)
            if (!exists || exists[u]) scores[u] /= norm_factor;
    }

    double mx = 1.0;
    if (normalized) {
        mx = scores[0];
        for (uint64_t u = 0; u < z; ++u)
            if ((!exists || exists[u]) && scores[u] > mx) mx = scores[u];
    }
    if (iterations_out) *iterations_out = it;
    if (maximum_out) *maximum_out = mx;
    free(deg); free(cnt); free(pr);
    return 0;
}

int oracle_pagerank(uint64_t z, const uint8_t *exists, const uint64_t *in_ptr,
                    const uint64_t *in_src, const double *in_w, double damp, double tol,
                    int norm, uint64_t max_iterations, int normalized, int sink_mode,
                    int num_threads, double *scores, uint64_t *iterations_out,
                    double *maximum_out) {
    return oracle_pagerank_tv(z, exists, in_ptr, in_src, in_w, damp, tol, norm, max_iterations,
                              normalized, sink_mode, num_threads, NULL, scores, iterations_out,
                              maximum_out);
}

/* Centrality::ranking order (Centrality.cpp:29-41): existing nodes, score descending, ties by
 * id ascending.  Writes the order into ids_out (length = number of existing nodes). */
static const double *g_rank_scores;
static int rank_cmp(const void *pa, const void *pb) {
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    double sa = g_rank_scores[a], sb = g_rank_scores[b];
    if (sa == sb) return (a < b) ? -1 : (a > b);
    return (sa > sb) ? -1 : 1;
}

uint64_t oracle_ranking(uint64_t z, const uint8_t *exists, const double *scores,
                        uint64_t *ids_out) {
    uint64_t k = 0;
    for (uint64_t u = 0; u < z; ++u)
        if (!exists || exists[u]) ids_out[k++] = u;
    g_rank_scores = scores;
    qsort(ids_out, k, sizeof(uint64_t), rank_cmp);
    return k;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
