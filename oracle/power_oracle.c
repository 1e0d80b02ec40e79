/*
 * power_oracle.c -- CPU restatement of NetworKit's EigenvectorCentrality and KatzCentrality
 * (TEST INFRASTRUCTURE ONLY).
 *
 * This file belongs to the oracle: tests load it (through ctypes) to check the GPU library
 * against.  The product path never links, loads or calls it.
 *
 * What it restates (paths in kit-parco/networkit v9.0):
 *   - EigenvectorCentrality::run          networkit/cpp/centrality/EigenvectorCentrality.cpp:21-66
 *       values and scoreData start at 1 in every slot (:22-23); each iteration pulls
 *       values[u] = sum_{in-edges (u, v, w)} w * scoreData[v] over existing nodes (:35-40),
 *       length = sqrt(sum values^2) over existing nodes (:43-46), values /= length (:49-51),
 *       swap (:53), until Aux::NumericTools::equal(length, oldLength, tol) (:54,
 *       NumericTools.cpp:16-18: x <= y + e and x >= y - e); then fabs of every existing score
 *       when scoreData[0] < 0 (:57-62).
 *   - KatzCentrality::run                 networkit/cpp/centrality/KatzCentrality.cpp:28-76
 *       values start at 1 in every slot (:29-31); each iteration sets
 *       values[u] = sum_{in-neighbours (v, w)} w * alpha * (1 + scoreData[v]) + beta over existing
 *       nodes (:45-59, the default InEdges direction), length = sqrt(sum values^2) (:62-65),
 *       scoreData = values (:67), until equal(length, oldLength, tol) (:69); then
 *       scoreData[u] /= length over existing nodes (:71-73).
 *   - the default alpha                   KatzCentrality.cpp:15-17 and GraphTools.cpp:14-27:
 *       1 / (1 + the largest unweighted out-degree over all slots)
 *   - in-edge semantics                   Graph.hpp:1918-1938 (directed: inEdges[u]; undirected:
 *                                          the adjacency of u, a self-loop listed once)
 *
 * Layout: in_ptr[z+1], in_src[m] (uint64), in_w[m] (double, NULL when unweighted: every weight
 * is NetworKit's defaultEdgeWeight 1.0), exists[z] (NULL when every slot exists).
 *
 * Parity pinning (tests/test_oracle.py): the reference's known answers (8-node graph, +-1e-4,
 * CentralityGTest.cpp:541-578; wiki-Vote Katz top-1 = node 699 with alpha 5e-4, :216-225),
 * dense eigendecompositions of the adjacency matrix, the sparse direct solve of
 * (I - alpha A) x = alpha A 1 + beta 1, and invariants (uniform scores on regular graphs).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int node_exists(const uint8_t *exists, uint64_t u) { return !exists || exists[u]; }

/* values = (scale * A) applied as in the reference loops; returns sqrt(sum values^2). */
static double pull(uint64_t z, const uint8_t *exists, const uint64_t *in_ptr,
                   const uint64_t *in_src, const double *in_w, const double *x, double *values,
                   int katz, double alpha, double beta) {
    double ss = 0.0;
#pragma omp parallel for schedule(guided) reduction(+ : ss)
    for (int64_t ui = 0; ui < (int64_t)z; ++ui) {
        const uint64_t u = (uint64_t)ui;
        if (!node_exists(exists, u)) continue;
        double acc = 0.0;
        for (uint64_t e = in_ptr[u]; e < in_ptr[u + 1]; ++e) {
            const double w = in_w ? in_w[e] : 1.0;
            if (katz) acc += w * alpha * (1.0 + x[in_src[e]]);
            else acc += w * x[in_src[e]];
        }
        if (katz) acc += beta;
        values[u] = acc;
        ss += acc * acc;
    }
    return sqrt(ss);
}

static int close_enough(double x, double y, double tol) { return x <= y + tol && x >= y - tol; }

/* EigenvectorCentrality::run.  scores[z] receives the result; *iterations the loop count.
 * max_iterations caps the loop (the reference has no cap).  Returns 0, or -1 on bad input. */
int oracle_eigenvector(uint64_t z, const uint8_t *exists, const uint64_t *in_ptr,
                       const uint64_t *in_src, const double *in_w, double tol,
                       uint64_t max_iterations, double *scores, uint64_t *iterations) {
    if (!scores || (z && (!in_ptr))) return -1;
    double *values = (double *)malloc(sizeof(double) * (z ? z : 1));
    if (!values) return -1;
    for (uint64_t u = 0; u < z; ++u) values[u] = scores[u] = 1.0;
    double *cur = scores, *nxt = values;
    double length = 0.0, old;
    uint64_t it = 0;
    do {
        old = length;
        length = pull(z, exists, in_ptr, in_src, in_w, cur, nxt, 0, 0.0, 0.0);
        for (uint64_t u = 0; u < z; ++u)
            if (node_exists(exists, u)) nxt[u] /= length;
        double *t = cur;
        cur = nxt;
        nxt = t;
        ++it;
    } while (!close_enough(length, old, tol) && it < max_iterations);
    if (cur != scores) memcpy(scores, cur, sizeof(double) * z);
    if (z && scores[0] < 0)
        for (uint64_t u = 0; u < z; ++u)
            if (node_exists(exists, u)) scores[u] = fabs(scores[u]);
    free(values);
    *iterations = it;
    return 0;
}

/* KatzCentrality::run with the given alpha (pass the default from oracle_katz_default_alpha).
 * scores[z] receives the result; *iterations the loop count. */
int oracle_katz(uint64_t z, const uint8_t *exists, const uint64_t *in_ptr,
                const uint64_t *in_src, const double *in_w, double alpha, double beta,
                double tol, uint64_t max_iterations, double *scores, uint64_t *iterations) {
    if (!scores || (z && (!in_ptr)) || alpha < 0) return -1;
    double *values = (double *)malloc(sizeof(double) * (z ? z : 1));
    if (!values) return -1;
    for (uint64_t u = 0; u < z; ++u) values[u] = scores[u] = 1.0;
    double length = 0.0, old;
    uint64_t it = 0;
    do {
        old = length;
        length = pull(z, exists, in_ptr, in_src, in_w, scores, values, 1, alpha, beta);
        memcpy(scores, values, sizeof(double) * z);
        ++it;
    } while (!close_enough(length, old, tol) && it < max_iterations);
    for (uint64_t u = 0; u < z; ++u)
        if (node_exists(exists, u)) scores[u] /= length;
    free(values);
    *iterations = it;
    return 0;
}

/* 1 / (1 + max unweighted out-degree over all slots): out-degrees counted from the in-lists
 * (an undirected adjacency counts every neighbour, a self-loop once). */
double oracle_katz_default_alpha(uint64_t z, const uint64_t *in_ptr, const uint64_t *in_src,
                                 int directed) {
    uint64_t *cnt = (uint64_t *)calloc(z ? z : 1, sizeof(uint64_t));
    if (!cnt) return -1.0;
    uint64_t best = 0;
    if (directed) {
        for (uint64_t u = 0; u < z; ++u)
            for (uint64_t e = in_ptr[u]; e < in_ptr[u + 1]; ++e) cnt[in_src[e]]++;
    } else {
        for (uint64_t u = 0; u < z; ++u) cnt[u] = in_ptr[u + 1] - in_ptr[u];
    }
    for (uint64_t u = 0; u < z; ++u)
        if (cnt[u] > best) best = cnt[u];
    free(cnt);
    return 1.0 / (1.0 + (double)best);
}
