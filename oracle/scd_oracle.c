/*
 * scd_oracle.c -- CPU restatement of NetworKit's ApproximatePageRank and of PageRankNibble's
 * sweep (TEST INFRASTRUCTURE ONLY).
 *
 * This file belongs to the oracle: tests load it (through ctypes) to check the GPU library
 * against.  The product path never links, loads or calls it.
 *
 * What it restates (paths in kit-parco/networkit v9.0):
 *   - ApproximatePageRank::run(seeds)     networkit/cpp/scd/ApproximatePageRank.cpp:18-61
 *       every seed starts with residual 1/|seeds| (the seed set is a std::set: ids ascending,
 *       no repeats) and enters a FIFO queue; push(u): res = r[u], vol = weightedDegree(u,
 *       countSelfLoopsTwice = true); for every neighbour v of u (Graph::forNeighborsOf, the
 *       out-list in storage order) mass = (1 - alpha) res w / (2 vol); v is queued when
 *       r[v] < vol_v eps and r[v] + mass >= eps vol_v; r[v] += mass; then p[u] += alpha res and
 *       r[u] = (1 - alpha) res / 2 (an assignment: mass that reached u through a self-loop in
 *       this push is dropped, as in the reference); u is queued again when r[u] / vol >= eps.
 *       The result holds every node the run touched (the keys of prRes), p possibly 0.
 *   - PageRankNibble::bestSweepSet        networkit/cpp/scd/PageRankNibble.cpp:28-89
 *       score = p / weightedDegree(u, true), sorted descending (the reference's sort is not
 *       stable; here equal scores keep ascending ids, one of the orders it may produce); the
 *       sweep adds the nodes in that order, volume += sum of the node's neighbour weights,
 *       cut += w for a neighbour outside the set so far (the node itself included) and -= w
 *       for one inside; conductance = cut / min(volume, 2 totalEdgeWeight - volume); the best
 *       prefix is the first with the smallest conductance among prefixes smaller than n.
 *
 * Layout: out-lists out_ptr[z+1], out_dst[m] (uint64), out_w[m] (double, NULL = unit weights).
 *
 * Parity pinning (tests/test_oracle.py): the ACL identity p + pr(r) = pr(seeds) against a
 * sparse direct solve of the lazy personalised PageRank on small graphs, the stopping
 * invariant r[u] < eps vol(u) everywhere, mass conservation, and the sweep against a brute
 * force over every prefix with set-based cut and volume.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double weighted_degree2(const uint64_t *out_ptr, const uint64_t *out_dst,
                               const double *out_w, uint64_t u) {
    /* Graph::weightedDegree(u, countSelfLoopsTwice = true), Graph.cpp:427-455 */
    double s = 0.0;
    for (uint64_t e = out_ptr[u]; e < out_ptr[u + 1]; ++e) {
        const double w = out_w ? out_w[e] : 1.0;
        s += out_dst[e] == u ? 2.0 * w : w;
    }
    return s;
}

/* The reference's weightedDegree is O(1) on unweighted graphs without self-loops and a scan of
 * the neighbours otherwise (Graph.cpp:426-453); here each node's value is computed once, on
 * first use, and kept, so the restatement costs what the reference costs on its usual inputs. */
struct volumes {
    double *v;
    uint8_t *known;
};

static double volume_of(struct volumes *vs, const uint64_t *out_ptr, const uint64_t *out_dst,
                        const double *out_w, uint64_t u) {
    if (!vs->known[u]) {
        vs->v[u] = weighted_degree2(out_ptr, out_dst, out_w, u);
        vs->known[u] = 1;
    }
    return vs->v[u];
}

struct queue {
    uint64_t *a;
    uint64_t cap, head, size;
};

static int q_push(struct queue *q, uint64_t v) {
    if (q->size == q->cap) {
        uint64_t nc = q->cap ? 2 * q->cap : 1024;
        uint64_t *na = (uint64_t *)malloc(nc * sizeof(uint64_t));
        if (!na) return -1;
        for (uint64_t i = 0; i < q->size; ++i) na[i] = q->a[(q->head + i) % q->cap];
        free(q->a);
        q->a = na;
        q->cap = nc;
        q->head = 0;
    }
    q->a[(q->head + q->size) % q->cap] = v;
    q->size++;
    return 0;
}

static uint64_t q_pop(struct queue *q) {
    uint64_t v = q->a[q->head];
    q->head = (q->head + 1) % q->cap;
    q->size--;
    return v;
}

/* p[z], r[z] receive the pairs of prRes (0 where untouched); touched[z] marks its keys.
 * seeds are sorted ascending without repeats by the caller (std::set).  Returns the number of
 * pushes, or -1 on allocation failure. */
int64_t oracle_apr(uint64_t z, const uint64_t *out_ptr, const uint64_t *out_dst,
                   const double *out_w, double alpha, double eps, const uint64_t *seeds,
                   uint64_t nseeds, double *p, double *r, uint8_t *touched) {
    memset(p, 0, z * sizeof(double));
    memset(r, 0, z * sizeof(double));
    memset(touched, 0, z);
    struct queue q = {NULL, 0, 0, 0};
    struct volumes vs = {(double *)malloc(z * sizeof(double)), (uint8_t *)calloc(z, 1)};
    if (!vs.v || !vs.known) {
        free(vs.v);
        free(vs.known);
        return -1;
    }
    const double init = 1.0 / (double)nseeds;
    for (uint64_t i = 0; i < nseeds; ++i) {
        const uint64_t s = seeds[i];
        p[s] = 0.0;
        r[s] = init;
        touched[s] = 1;
        if (q_push(&q, s)) goto fail;
    }
    int64_t pushes = 0;
    while (q.size) {
        const uint64_t u = q_pop(&q);
        touched[u] = 1;
        const double res = r[u];
        const double volume = volume_of(&vs, out_ptr, out_dst, out_w, u);
        for (uint64_t e = out_ptr[u]; e < out_ptr[u + 1]; ++e) {
            const uint64_t v = out_dst[e];
            const double w = out_w ? out_w[e] : 1.0;
            const double mass = (1.0 - alpha) * res * w / (2.0 * volume);
            const double volv = volume_of(&vs, out_ptr, out_dst, out_w, v);
            touched[v] = 1;
            if (r[v] < volv * eps && (r[v] + mass) >= eps * volv)
                if (q_push(&q, v)) goto fail;
            r[v] += mass;
        }
        p[u] = p[u] + alpha * res;
        r[u] = (1.0 - alpha) * res / 2;
        if ((r[u] / volume) >= eps)
            if (q_push(&q, u)) goto fail;
        ++pushes;
    }
    free(q.a);
    free(vs.v);
    free(vs.known);
    return pushes;
fail:
    free(q.a);
    free(vs.v);
    free(vs.known);
    return -1;
}

/* bestSweepSet on the k nodes ids[k] with scores p[k] (already divided by the degree or not:
 * the caller passes score = p / weightedDegree(u, true)); the order is written to order[k]
 * and the best prefix length is returned (0 when none qualifies).  n = numberOfNodes,
 * total_volume = 2 totalEdgeWeight. */
static const double *g_score;
static const uint64_t *g_ids;
static int sweep_cmp(const void *a, const void *b) {
    const uint64_t i = *(const uint64_t *)a, j = *(const uint64_t *)b;
    const double si = g_score[i], sj = g_score[j];
    if (si > sj) return -1;
    if (si < sj) return 1;
    return g_ids[i] < g_ids[j] ? -1 : (g_ids[i] > g_ids[j] ? 1 : 0);
}

uint64_t oracle_sweep(uint64_t z, const uint64_t *out_ptr, const uint64_t *out_dst,
                      const double *out_w, uint64_t n, double total_volume, uint64_t k,
                      const uint64_t *ids, const double *score, uint64_t *order,
                      double *best_cond) {
    uint64_t *idx = (uint64_t *)malloc((k + 1) * sizeof(uint64_t));
    uint8_t *in = (uint8_t *)calloc(z + 1, 1);
    if (!idx || !in) {
        free(idx);
        free(in);
        return 0;
    }
    for (uint64_t i = 0; i < k; ++i) idx[i] = i;
    g_score = score;
    g_ids = ids;
    qsort(idx, k, sizeof(uint64_t), sweep_cmp);
    double best = INFINITY, cut = 0.0, volume = 0.0;
    uint64_t best_len = 0;
    for (uint64_t i = 0; i < k; ++i) {
        const uint64_t v = ids[idx[i]];
        order[i] = v;
        double wdeg = 0.0;
        for (uint64_t e = out_ptr[v]; e < out_ptr[v + 1]; ++e) {
            const double w = out_w ? out_w[e] : 1.0;
            wdeg += w;
            if (!in[out_dst[e]]) cut += w;
            else cut -= w;
        }
        volume += wdeg;
        in[v] = 1;
        const double rest = total_volume - volume;
        const double cond = cut / (volume < rest ? volume : rest);
        if (cond < best && (i + 1) < n) {
            best = cond;
            best_len = i + 1;
        }
    }
    if (best_cond) *best_cond = best;
    free(idx);
    free(in);
    return best_len;
}
