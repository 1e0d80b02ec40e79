/*
 * algebraic_oracle.c -- CPU restatement of NetworKit's AlgebraicPageRank and of the CSR
 * matrix-vector product it runs on (TEST INFRASTRUCTURE ONLY).
 *
 * This file belongs to the oracle: tests load it (through ctypes) to check the GPU library
 * against.  The product path never links, loads or calls it.
 *
 * What it restates (paths in kit-parco/networkit v9.0):
 *   - CSRGeneralMatrix<double>(nRows, nCols, triplets)   include/networkit/algebraic/
 *       CSRGeneralMatrix.hpp:151-182: a counting sort of the triplets by row that keeps their
 *       order inside a row; duplicates stay separate entries
 *   - CSRGeneralMatrix::adjacencyMatrix                   CSRGeneralMatrix.hpp:889-900: one
 *       triplet (u, v, w) per edge of Graph::forEdges (u ascending, each out-list in insertion
 *       order; undirected graphs visit an edge at its endpoint u >= v, Graph.hpp:1895-1915,
 *       1941-1945) and (v, u, w) after it when undirected and u != v; dimension z
 *   - GraphBLAS::rowReduce                                GraphBLAS.hpp:278-291 (row sums)
 *   - CSRGeneralMatrix::operator*(const Vector&)          CSRGeneralMatrix.hpp:563-578:
 *       y[i] = sum over row i in storage order of nonZeros[k] * x[columnIdx[k]]
 *   - AlgebraicPageRank(G, damp, tol)                      include/networkit/algebraic/
 *       algorithms/AlgebraicPageRank.hpp:35-54: M = the triplets (j, i, damp * A(i, j) *
 *       (1 / rowsum_i)) taken row by row in A's storage order
 *   - AlgebraicPageRank::run                               AlgebraicPageRank.hpp:94-119:
 *       rank = 1/n (n = rows = z), do { last = rank; rank = M rank + (1 - damp)/n } while
 *       ||rank - last||_2 > tol; then rank / sum(rank) over every slot
 *
 * The reference loop has no iteration cap; max_it (UINT64_MAX for none) bounds it here.
 *
 * Parity pinning (tests/test_oracle.py): the reference's known answers
 * (AlgebraicPageRankGTest.cpp:21-94: wiki-Vote top-1 node 699, score 0.00432 +- 1e-3; the
 * 8-node weighted graph +- 1e-4), the CSR product against scipy.sparse, the CSR build against
 * a dense matrix, and the fixed point against a sparse direct solve of (I - M) x = (1-d)/n 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* CSRGeneralMatrix(nRows, nCols, triplets): counting sort by row, order kept inside a row. */
static int csr_from_triplets(uint64_t rows, uint64_t nnz, const uint64_t *tr, const uint64_t *tc,
                             const double *tv, uint64_t *row_idx, uint64_t *col_idx,
                             double *values) {
    uint64_t *next = (uint64_t *)calloc(rows + 1, sizeof(uint64_t));
    if (!next) return -1;
    for (uint64_t i = 0; i < nnz; ++i) next[tr[i] + 1]++;
    for (uint64_t r = 0; r < rows; ++r) next[r + 1] += next[r];
    memcpy(row_idx, next, (rows + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < nnz; ++i) {
        uint64_t d = next[tr[i]]++;
        col_idx[d] = tc[i];
        values[d] = tv[i];
    }
    free(next);
    return 0;
}

/* CSRGeneralMatrix::operator*(const Vector&): rows in parallel, as the reference's
 * `omp parallel for`; each row sums in storage order. */
void oracle_csr_spmv(uint64_t rows, const uint64_t *row_idx, const uint64_t *col_idx,
                     const double *values, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)rows; ++i) {
        double sum = 0.0;
        for (uint64_t k = row_idx[i]; k < row_idx[i + 1]; ++k) sum += values[k] * x[col_idx[k]];
        y[i] = sum;
    }
}

/* The matrix M of AlgebraicPageRank(G, damp) from the edges of G in Graph::forEdges order
 * (eu[k], ev[k], ew[k]; ew NULL = unit weights).  Writes M in CSR form: row_idx[z+1],
 * col_idx[nnz], values[nnz] with nnz = edges (directed) or 2 * edges - self-loops
 * (undirected); the caller sizes them with oracle_algebraic_nnz. */
uint64_t oracle_algebraic_nnz(uint64_t edges, const uint64_t *eu, const uint64_t *ev,
                              int directed) {
    uint64_t nnz = edges;
    if (!directed)
        for (uint64_t k = 0; k < edges; ++k) nnz += eu[k] != ev[k];
    return nnz;
}

int oracle_algebraic_matrix(uint64_t z, uint64_t edges, const uint64_t *eu, const uint64_t *ev,
                            const double *ew, int directed, double damp, uint64_t *m_row,
                            uint64_t *m_col, double *m_val) {
    const uint64_t nnz = oracle_algebraic_nnz(edges, eu, ev, directed);
    uint64_t *tr = (uint64_t *)malloc((nnz + 1) * sizeof(uint64_t));
    uint64_t *tc = (uint64_t *)malloc((nnz + 1) * sizeof(uint64_t));
    double *tv = (double *)malloc((nnz + 1) * sizeof(double));
    uint64_t *a_row = (uint64_t *)malloc((z + 1) * sizeof(uint64_t));
    uint64_t *a_col = (uint64_t *)malloc((nnz + 1) * sizeof(uint64_t));
    double *a_val = (double *)malloc((nnz + 1) * sizeof(double));
    double *inv = (double *)malloc((z + 1) * sizeof(double));
    int rc = -1;
    if (!tr || !tc || !tv || !a_row || !a_col || !a_val || !inv) goto out;
    /* adjacencyMatrix: triplets in forEdges order */
    uint64_t t = 0;
    for (uint64_t k = 0; k < edges; ++k) {
        const double w = ew ? ew[k] : 1.0;
        tr[t] = eu[k]; tc[t] = ev[k]; tv[t] = w; ++t;
        if (!directed && eu[k] != ev[k]) {
            tr[t] = ev[k]; tc[t] = eu[k]; tv[t] = w; ++t;
        }
    }
    if (csr_from_triplets(z, nnz, tr, tc, tv, a_row, a_col, a_val)) goto out;
    /* rowReduce, then 1 / row sum (inf for an empty row, never used) */
    for (uint64_t i = 0; i < z; ++i) {
        double s = 0.0;
        for (uint64_t k = a_row[i]; k < a_row[i + 1]; ++k) s += a_val[k];
        inv[i] = 1.0 / s;
    }
    /* M triplets (j, i, damp * value * inv[i]) in A's row order */
    t = 0;
    for (uint64_t i = 0; i < z; ++i)
        for (uint64_t k = a_row[i]; k < a_row[i + 1]; ++k) {
            tr[t] = a_col[k]; tc[t] = i; tv[t] = damp * a_val[k] * inv[i]; ++t;
        }
    rc = csr_from_triplets(z, nnz, tr, tc, tv, m_row, m_col, m_val);
out:
    free(tr); free(tc); free(tv); free(a_row); free(a_col); free(a_val); free(inv);
    return rc;
}

/* AlgebraicPageRank::run on M (CSR, n rows).  scores[n] receives the sum-normalised result. */
int oracle_algebraic_pagerank(uint64_t n, const uint64_t *m_row, const uint64_t *m_col,
                              const double *m_val, double damp, double tol, uint64_t max_it,
                              double *scores, uint64_t *iterations) {
    double *last = (double *)malloc((n + 1) * sizeof(double));
    if (!last) return -1;
    const double tele = (1.0 - damp) / (double)n;
    for (uint64_t i = 0; i < n; ++i) scores[i] = 1.0 / (double)n;
    uint64_t it = 0;
    double step;
    do {
        memcpy(last, scores, n * sizeof(double));
        oracle_csr_spmv(n, m_row, m_col, m_val, last, scores);
        step = 0.0;
#pragma omp parallel for reduction(+ : step)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            scores[i] += tele;
            const double d = scores[i] - last[i];
            step += d * d;
        }
        ++it;
    } while (sqrt(step) > tol && it < max_it);
    double sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (int64_t i = 0; i < (int64_t)n; ++i) sum += scores[i];
#pragma omp parallel for
    for (int64_t i = 0; i < (int64_t)n; ++i) scores[i] /= sum;
    *iterations = it;
    free(last);
    return 0;
}
