/*
 * nkpr.h -- C ABI of the MI355X PageRank library (libnkpr.so).
 *
 * This is the boundary a NetworKit build would bind instead of running the OpenMP loop in
 * networkit/cpp/centrality/PageRank.cpp.  Every entry point takes plain pointers and sizes,
 * returns an int status (NKPR_OK = 0) and never throws; the message of the last failure on
 * the calling thread is available from nkpr_last_error().
 *
 * Which reference interface each entry point replaces (paths in kit-parco/networkit v9.0):
 *   nkpr_graph_create    the Graph state PageRank::run reads: in-lists (Graph.hpp:1918-1938),
 *                        exists[] (Graph.hpp:1802-1809), weightedDegree (Graph.cpp:427-455)
 *   nkpr_run             PageRank::run (PageRank.cpp:20-104) incl. numberOfIterations
 *                        (PageRank.hpp:58-61) and maximum() (PageRank.cpp:95-108)
 *   nkpr_ranking         Centrality::ranking (Centrality.cpp:29-41)
 *   nkpr_ppr_batch       no counterpart: PageRank::run once per teleport distribution
 *   nkpr_adjoint         no counterpart: gradients of the scores (differentiable PageRank)
 *   nkpr_graph_create_batch / nkpr_run_batch  no counterpart: PageRank::run once per graph
 *   nkpr_graph_set_weights  no counterpart: a new Graph with the same structure
 *   nkpr_params_default  the constructor defaults PageRank(G, 0.85, 1e-8, false), the field
 *                        defaults maxIterations = UINT64_MAX, norm = L2Norm (PageRank.hpp:40-67)
 *   nkpr_power_run       EigenvectorCentrality::run (EigenvectorCentrality.cpp:21-66) and
 *                        KatzCentrality::run (KatzCentrality.cpp:28-76)
 *   nkpr_power_params_default  the constructor defaults EigenvectorCentrality(G, 1e-8)
 *                        (EigenvectorCentrality.hpp:31) and KatzCentrality(G, 0, 0.1, 1e-8)
 *                        (KatzCentrality.hpp:37)
 *
 * Device placement: a graph lives on the HIP device that was current when it was created and
 * owns one HIP stream (default priority); all of its work is ordered on that stream.  Calls
 * on one graph handle must not overlap (the reference does not support concurrent run() on
 * one object either).  Distinct handles may be used from distinct threads.
 */
#ifndef NKPR_H
#define NKPR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NKPR_ABI_VERSION 3

/* status codes */
#define NKPR_OK 0
#define NKPR_ERR_INVALID 1     /* bad argument or unsupported input                      */
#define NKPR_ERR_HIP 2         /* a HIP runtime call failed (message has the HIP error)   */
#define NKPR_ERR_NOMEM 3       /* device or host allocation failed                        */
#define NKPR_ERR_INTERRUPTED 4 /* *interrupt became nonzero (Aux::SignalHandler analogue) */
#define NKPR_ERR_NODEVICE 5    /* no usable gfx950 device                                 */

/* PageRank::Norm (PageRank.hpp:31): L1Norm = 0, L2Norm = 1 */
#define NKPR_L1 0
#define NKPR_L2 1

/* sink handling.  NONE is the reference behaviour: rank sent to a node without out-edges is
 * lost every sweep and only the final normalisation rescales it (PageRank.cpp:56-92).
 * DISTRIBUTE adds damp * (sum of sink scores) / n to every node each sweep (Google matrix). */
#define NKPR_SINK_NONE 0
#define NKPR_SINK_DISTRIBUTE 1

/* arithmetic of the sweeps.
 * FP64  every sweep in FP64 (folded FP64 edge coefficients); with NKPR_STOP_STEP it follows
 *       the reference loop exactly (same stopping rule and iteration count).
 * MIXED FP32 sweeps until the step stalls, then FP64 residual sweeps with FP32 correction
 *       solves (iterative refinement) until the certified L1 error of the sum-normalised
 *       scores is <= tol.  Always uses NKPR_STOP_L1_ERROR.
 * FP32  every sweep in FP32 (stops at the FP32 noise floor if tol is below it).  With
 *       NKPR_STOP_L1_ERROR it stops when the formula of that rule, evaluated on the FP32
 *       iterates, is <= tol; that is not a certificate (FP32 rounding is not part of it), and
 *       the reported l1_error_bound is +inf. */
#define NKPR_FP64 0
#define NKPR_MIXED 1
#define NKPR_FP32 2

/* stopping rule.
 * STEP      the reference rule: stop when ||x' - x|| <= tol in `norm` (PageRank.cpp:34-51)
 *           or the iteration count reaches max_iterations.
 * L1_ERROR  stop when the certified bound on the L1 error of the sum-normalised scores,
 *           2 * d/(1-d) * ||x' - x||_1 / sum(x'), is <= tol. */
#define NKPR_STOP_STEP 0
#define NKPR_STOP_L1_ERROR 1

typedef struct nkpr_graph nkpr_graph;

/* In-lists of a graph in compressed form (CSC of the pull operator).  Row u lists the sources
 * of u's in-edges: for a directed graph the inEdges[u] of NetworKit, for an undirected graph
 * its adjacency outEdges[u] (both endpoints list an edge, a self-loop is listed once). */
typedef struct {
    uint64_t z;               /* node-id bound (Graph::upperNodeIdBound); z < 2^31        */
    const uint64_t *colptr;   /* [z+1] offsets into src/weight; colptr[z] = m < 2^32     */
    const int32_t *src;       /* [m] source node of each in-edge                           */
    const float *weight;      /* [m] edge weights, or NULL: unweighted (every weight 1.0)  */
    const double *weight64;   /* [m] optional exact FP64 weights (NetworKit edgeweight); when
                                 given they define the FP64 coefficients, `weight` must then
                                 hold the same values rounded to FP32                      */
    const double *out_wdeg;   /* [z] weighted out-degree per node, or NULL: summed from the
                                 lists (Graph::weightedDegree(u) semantics)               */
    const uint8_t *exists;    /* [z] 1 if the slot holds a node, or NULL: all exist        */
    int32_t device_ptrs;      /* 1: the arrays are device pointers on the current device  */
} nkpr_csc;

typedef struct {
    uint64_t z;               /* node-id bound                                             */
    uint64_t n;               /* existing nodes (Graph::numberOfNodes)                     */
    uint64_t m;               /* in-list entries                                           */
    uint64_t sinks;           /* existing nodes without out-edges                          */
    uint64_t max_in_degree;
    uint64_t work_items;      /* scheduled row blocks + split-row chunks per sweep         */
    uint64_t device_bytes;    /* device memory held by the handle                          */
    int32_t weighted;
    int32_t device;           /* HIP device ordinal                                        */
    double prep_ms;           /* wall time of nkpr_graph_create                            */
} nkpr_graph_info;

typedef struct {
    double damp;              /* damping factor d (default 0.85)                           */
    double tol;               /* tolerance (default 1e-8), meaning set by `stop`           */
    uint64_t max_iterations;  /* default UINT64_MAX                                        */
    int32_t norm;             /* NKPR_L1 / NKPR_L2 (default L2)                            */
    int32_t sink_handling;    /* NKPR_SINK_* (default NONE)                                */
    int32_t normalized;       /* PageRank(..., normalized) (default 0)                     */
    int32_t precision;        /* NKPR_FP64 / NKPR_MIXED / NKPR_FP32 (default FP64)         */
    int32_t stop;             /* NKPR_STOP_* (default STEP; MIXED always uses L1_ERROR)    */
    int32_t scores_on_device; /* 1: `scores` of nkpr_run is a device pointer               */
    int32_t timing;           /* 1: time every sweep kernel with HIP events                */
    int32_t reserved;
    const volatile int32_t *interrupt; /* optional; checked between batches of sweeps      */
    const double *teleport;   /* optional [z] teleport distribution t (personalised PageRank,
                                 not in the reference): node u receives (1-d) t[u] instead of
                                 (1-d)/n; host or device memory like `scores`.  NULL: the
                                 reference's uniform teleport.  nkpr_run only.            */
} nkpr_params;

typedef struct {
    uint64_t iterations;      /* PageRank::numberOfIterations(): all sweeps of the run     */
    uint64_t sweeps_fp32;
    uint64_t sweeps_fp64;
    uint64_t refinements;     /* FP64 residual sweeps followed by an FP32 correction solve */
    uint64_t sweeps_fp32_correction; /* FP32 sweeps of correction solves (part of sweeps_fp32) */
    double last_step;         /* ||x' - x|| of the last sweep in params->norm              */
    double last_step_l1;
    double l1_error_bound;    /* certified bound on the L1 error of the normalised scores
                                 (FP64 sweeps); +inf for NKPR_FP32, which certifies nothing */
    double maximum;           /* PageRank::maximum()                                       */
    double score_sum;         /* sum of the unnormalised scores before normalisation       */
    double ms_sweeps_fp32;    /* summed sweep-kernel time (params->timing = 1)             */
    double ms_sweeps_fp64;
    double ms_total;          /* wall time of nkpr_run                                     */
} nkpr_stats;

/* Fills the reference defaults. */
int nkpr_params_default(nkpr_params *p);

/* Uploads (or, with device_ptrs, copies on the device) and preprocesses a graph: relabels the
 * existing nodes by out-degree (sinks last), sorts every in-list by the new source id, folds
 * w / deg(src) into the edge coefficients and builds the sweep schedule. */
int nkpr_graph_create(const nkpr_csc *csc, nkpr_graph **out);
int nkpr_graph_destroy(nkpr_graph *g);
int nkpr_graph_get_info(const nkpr_graph *g, nkpr_graph_info *info);
/* The graph's HIP stream (hipStream_t), e.g. to order caller work or events with it. */
void *nkpr_graph_stream(const nkpr_graph *g);

/* PageRank::run.  scores[z] holds the start vector on entry (a fresh reference object starts
 * at 1/n everywhere; a re-run starts from the previous scores, PageRank.cpp:26) and the
 * scores on return; slots that hold no node are left untouched.  stats may be NULL. */
int nkpr_run(nkpr_graph *g, const nkpr_params *p, double *scores, nkpr_stats *stats);

/* New edge weights for a graph made by nkpr_graph_create with weights (same structure): the
 * arrays are laid out like nkpr_csc.weight / weight64 / out_wdeg (out_wdeg NULL: summed from
 * the weights); device or host memory per device_ptrs.  The folded coefficients w / deg are
 * recomputed in place; no counterpart in the reference, where a new Graph would be built. */
int nkpr_graph_set_weights(nkpr_graph *g, const float *weight, const double *weight64,
                           const double *out_wdeg, int32_t device_ptrs);

/* Adjoint of PageRank (no counterpart in the reference).  For the sum-normalised scores x of
 * nkpr_run (normalized = 0, NKPR_SINK_NONE, optional params->teleport t) and the gradient
 * grad_scores = dL/dx of a loss, returns
 *   grad_teleport[z] = dL/dt = (1-d) lambda,
 *   grad_weight[m]   = dL/dw for every in-edge of `csc` (laid out like csc->weight),
 * where lambda = d P^T lambda + (dL/dx - (dL/dx . x)) / score_sum.  scores and score_sum
 * (nkpr_stats.score_sum) come from the forward run; `csc` is the structure the graph was
 * created from, with its current weights.  params->precision: FP32 = restarted FP32 GMRES to a
 * relative 2-norm residual <= tol; FP64 / MIXED = FP64 residual sweeps with FP32 GMRES
 * corrections to a relative L1 residual <= tol.  z-vectors are host or device memory per
 * scores_on_device, grad_weight per csc->device_ptrs; either output may be NULL. */
int nkpr_adjoint(nkpr_graph *g, const nkpr_params *p, const nkpr_csc *csc,
                 const double *scores, double score_sum, const double *grad_scores,
                 double *grad_weight, double *grad_teleport, nkpr_stats *stats);

/* A batch of graphs as one handle (no counterpart in the reference).  `csc` is the
 * block-diagonal union of the graphs: graph q owns the slots [graph_offsets[q],
 * graph_offsets[q+1]) (graph_offsets[0] = 0, graph_offsets[num_graphs] = csc->z, host memory)
 * and no in-list entry may join two graphs. */
int nkpr_graph_create_batch(const nkpr_csc *csc, uint32_t num_graphs,
                            const uint64_t *graph_offsets, nkpr_graph **out);

/* PageRank::run on every graph of a batch handle at once.  Each graph follows the rules of
 * nkpr_run on its own: its own n, teleport (1-d)/n_q, sink mass, stopping rule and iteration
 * count (a graph that has converged is frozen), normalisation and maximum().  scores[z] is in
 * the union's slot order (start vectors on entry).  iterations[num_graphs] and
 * maximum[num_graphs] (host) are optional; stats->iterations counts the sweeps of the union. */
int nkpr_run_batch(nkpr_graph *g, const nkpr_params *p, double *scores, uint64_t *iterations,
                   double *maximum, nkpr_stats *stats);

/* Statistics of a batched run (nkpr_ppr_batch). */
typedef struct {
    uint64_t sweeps;          /* SpMM sweeps over all columns, all precisions              */
    uint64_t sweeps_fp32;
    uint64_t sweeps_fp64;
    uint64_t refinements;     /* FP64 residual sweeps followed by an FP32 correction solve */
    uint64_t max_iterations_col; /* largest per-column iteration count                     */
    double max_error_bound;   /* largest per-column certified L1 bound (MIXED / L1_ERROR;
                                 +inf for FP32) or last step in `norm` (STEP)              */
    double ms_sweeps_fp32;    /* summed SpMM kernel time (params->timing = 1)              */
    double ms_sweeps_fp64;
    double ms_total;
    uint32_t tile_columns;    /* column tile width chosen by the autotuner                 */
    uint32_t heavy_degree;    /* in-degree above which a row is split (autotuned)          */
} nkpr_batch_stats;

/* Batched personalised PageRank (no counterpart in the reference, which would run
 * PageRank::run once per teleport distribution, PageRank.cpp:20-104).  Column j uses the
 * teleport distribution uniform over seed_nodes[seed_offsets[j] .. seed_offsets[j+1]) (slot
 * ids; a repeated id counts repeatedly), or proportional to seed_weights there when given:
 * node u receives (1-d) t_j[u] instead of (1-d)/n.  Distributed sink mass stays uniform.
 * Every column follows nkpr_run's rules for `params` on its own: the stopping rule is applied
 * per column and a converged column is frozen; MIXED certifies every column.  All columns
 * advance together in one sparse-matrix x dense-matrix sweep per iteration.
 * scores: [k][z] (column j at scores + j*z), start vectors on entry, each column divided by
 * its sum on return (normalized must be 0); host or device memory per scores_on_device.
 * seed arrays are host memory.  col_iterations[k] and col_error_bound[k] are optional. */
int nkpr_ppr_batch(nkpr_graph *g, const nkpr_params *p, uint32_t k,
                   const uint64_t *seed_offsets, const int32_t *seed_nodes,
                   const double *seed_weights, double *scores, nkpr_batch_stats *stats,
                   uint64_t *col_iterations, double *col_error_bound);

/* Eigenvector and Katz centrality on a graph handle whose coefficients are the edge weights:
 * one created with out_wdeg[u] = 1 for every node (a handle made otherwise is rejected).  Both
 * are the power iterations of the reference, pulled over the handle's in-lists:
 *   NKPR_EIGENVECTOR  EigenvectorCentrality::run (EigenvectorCentrality.cpp:21-66):
 *                     y = A x, x = y / ||y||_2, until | ||y|| - previous ||y|| | <= tol;
 *                     |x| when the score of slot 0 is negative.
 *   NKPR_KATZ         KatzCentrality::run (KatzCentrality.cpp:28-76):
 *                     x = alpha A (1 + x) + beta, until | ||x|| - previous ||x|| | <= tol;
 *                     then x / ||x||_2.  alpha is used as given (the reference's default
 *                     1 / (1 + max out-degree) is the caller's to compute); for the reference's
 *                     edgeDirection OutEdges create the handle from the out-lists.
 * Both start from 1 in every slot; norms are over existing nodes; slots that hold no node are
 * set to 1 (as the reference leaves them).  The sweeps are FP64.  A Katz run whose norm stops
 * being finite (alpha above 1 / spectral radius) fails with NKPR_ERR_INVALID, where the
 * reference would loop forever.  stats: iterations (sweeps), last_step = the final change of
 * the norm, score_sum = the final norm, ms_total. */
#define NKPR_EIGENVECTOR 0
#define NKPR_KATZ 1
typedef struct {
    int32_t kind;             /* NKPR_EIGENVECTOR or NKPR_KATZ                             */
    int32_t scores_on_device; /* 1: `scores` is a device pointer                           */
    double alpha;             /* Katz: damping of the product (>= 0)                       */
    double beta;              /* Katz: constant added to every node                        */
    double tol;               /* stop when the norm changes by at most tol (default 1e-8)  */
    uint64_t max_iterations;  /* not in the reference, which loops until the test holds;
                                 default UINT64_MAX                                        */
    const volatile int32_t *interrupt; /* optional; checked between batches of sweeps      */
} nkpr_power_params;

/* Fills kind EIGENVECTOR, alpha 0, beta 0.1, tol 1e-8, max_iterations UINT64_MAX. */
int nkpr_power_params_default(nkpr_power_params *p);
/* scores[z]: output only (both reference loops start from 1, with no warm start). */
int nkpr_power_run(nkpr_graph *g, const nkpr_power_params *p, double *scores, nkpr_stats *stats);

/* ---- CSR matrices and AlgebraicPageRank (NetworKit's algebraic module) -------------------
 * A CSRGeneralMatrix<double> in NetworKit's own layout (include/networkit/algebraic/
 * CSRGeneralMatrix.hpp: rowIdx, columnIdx, nonZeros), every array in device memory of the
 * current device: row_idx[rows + 1] rises from 0 to nnz, col_idx[nnz] < cols, values[nnz].
 * Entries of a row are summed in storage order groups; duplicates stay separate entries. */
typedef struct {
    uint64_t rows;
    uint64_t cols;
    uint64_t nnz;
    const uint64_t *row_idx;
    const uint64_t *col_idx;
    const double *values;
} nkpr_csr;

/* Replaces CSRGeneralMatrix::operator*(const Vector&) (CSRGeneralMatrix.hpp:563-578):
 * y[i] = sum over row i of values[k] * x[col_idx[k]].  x[cols] and y[rows] are distinct device
 * vectors; rows < 2^32 - 1.  Enqueued on `stream` (a hipStream_t; NULL is the null stream);
 * returns without waiting.  Its scratch comes from the stream's memory pool.  The matrix
 * structure is taken as given (see nkpr_algebraic_pagerank for a run that checks it). */
int nkpr_csr_spmv(const nkpr_csr *a, const double *x, double *y, void *stream);

/* One iteration of AlgebraicPageRank::run (AlgebraicPageRank.hpp:97-101) on a square matrix
 * M: y = M x + shift and *step_sq = ||y - x||_2^2 (FP64, a device pointer; summed in an order
 * fixed by the matrix).  Stream-ordered like nkpr_csr_spmv.  Like nkpr_csr_spmv it takes its
 * scratch (the list of rows longer than 32 x the lanes per row, their pieces of 4096 entries,
 * block sums: about 8 bytes per 4096 entries) from the stream's memory pool for the call. */
int nkpr_csr_pagerank_step(const nkpr_csr *m, double shift, const double *x, double *y,
                           double *step_sq, void *stream);

/* No counterpart in the reference: nkpr_csr_pagerank_step for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results; its scratch lives in a 48.1 MiB device array of the
 * library (held by every GPU the library runs on until the process ends) instead of the
 * stream's pool, and it calls no allocating routine.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads) race and must be
 * ordered by the caller.  A matrix that could have more than 2^21 rows longer than 32 x the
 * lanes per row (at most nnz / (32 lanes + 1) of them; 4 lanes up to an average row length of
 * 6, 8 up to 12, 16 up to 24, 32 up to 48, else 64) or 2^22 pieces of 4096 entries is rejected
 * with NKPR_ERR_INVALID before any launch: up to about 2.7e8 entries are accepted at 4 lanes,
 * 1.07e9 at 16 (R-MAT scale 26, 1.06e9 entries, fits) and 4.3e9 at 64. */
int nkpr_csr_pagerank_step_noalloc(const nkpr_csr *m, double shift, const double *x, double *y,
                                   double *step_sq, void *stream);

/* Replaces AlgebraicPageRank<CSRMatrix>::run (AlgebraicPageRank.hpp:94-119) on the matrix M
 * its constructor builds (:35-54, M(j, i) = damp * A(i, j) / rowsum_i(A)): from 1/n in every
 * row, y = M x + (1 - damp)/n until ||y - x||_2 <= tol (at least one iteration), then
 * y / sum(y).  n = rows (every slot, as the reference; no exists[] handling).  The reference
 * has no iteration cap: pass UINT64_MAX for none.  scores[rows] is a host pointer, or a device
 * pointer when scores_on_device = 1.  The matrix is checked once (NKPR_ERR_INVALID when
 * row_idx or a column index is out of place). Blocks until done. */
int nkpr_algebraic_pagerank(const nkpr_csr *m, double damp, double tol, uint64_t max_iterations,
                            double *scores, int32_t scores_on_device, uint64_t *iterations);

/* AlgebraicPageRank as a handle over the matrix, for repeated runs.  nkpr_algebraic_create
 * checks M (device arrays, as nkpr_algebraic_pagerank takes them; square, rows < 2^31 - 1,
 * nnz < 2^32 - 1) and builds a PageRank handle whose in-lists are M's rows and whose
 * coefficients are M's values (the caller may free M's arrays afterwards).
 * nkpr_algebraic_run then replaces AlgebraicPageRank::run (AlgebraicPageRank.hpp:94-119) like
 * nkpr_algebraic_pagerank (same start, teleport (1 - damp) / n, loop condition
 * ||x' - x||_2 > tol, division by the sum; scores in row order, host or device), with the
 * PageRank sweeps of the handle (relabelled, source-sliced, binned FP64 from 2^26 entries):
 * each row is still the FP64 sum of its products, added in another order.  stats (nullable)
 * receives iterations, sweeps_fp64, ms_sweeps_fp64 (HIP events around every sweep on the
 * handle's stream), score_sum, last_step (the final ||x' - x||_2) and ms_total.  Blocks until
 * done. */
typedef struct nkpr_algebraic nkpr_algebraic;
int nkpr_algebraic_create(const nkpr_csr *m, nkpr_algebraic **out);
int nkpr_algebraic_run(nkpr_algebraic *h, double damp, double tol, uint64_t max_iterations,
                       double *scores, int32_t scores_on_device, uint64_t *iterations,
                       nkpr_stats *stats);
int nkpr_algebraic_destroy(nkpr_algebraic *h);

/* ---- NetworKit binary graph files (nkbg002) ------------------------------------------------
 * Host-side CSR lists of a graph in NetworKit's binary format (include/networkit/io/
 * NetworkitBinaryGraph.hpp:20-141).  Arrays are host memory owned by the struct (released by
 * nkpr_nkbg_free).  The pull lists are what PageRank reads: the in-lists of a directed graph,
 * the adjacency of an undirected one (u's entries with v <= u, then those with v > u, as the
 * reference's reader appends them); colptr/src/weight/exists plug into nkpr_csc as they are
 * (with device_ptrs = 0, weights as weight64). */
typedef struct {
    uint64_t nodes;          /* slots, exists[] marks the nodes                              */
    uint64_t edges;          /* Graph::numberOfEdges of the stored graph                      */
    uint64_t self_loops;
    int32_t directed;
    int32_t weighted;
    int32_t weight_format;   /* 0 none, 1 unsigned varint, 2 signed varint, 3 double, 4 float */
    int32_t reserved;
    uint64_t *colptr;        /* [nodes + 1] pull lists                                        */
    int32_t *src;            /* [colptr[nodes]]                                               */
    double *weight;          /* [colptr[nodes]], NULL when unweighted                         */
    uint64_t *out_ptr;       /* out-lists of a directed graph when requested, else NULL       */
    int32_t *out_dst;
    double *out_weight;
    uint8_t *exists;         /* [nodes]                                                       */
} nkpr_nkbg;

/* Replaces NetworkitBinaryReader::read (networkit/cpp/io/NetworkitBinaryReader.cpp:19-237)
 * without NetworKit's per-node vectors: decodes the file on host threads straight into the
 * lists above.  with_out_lists = 1 also fills out_ptr/out_dst/out_weight of a directed graph.
 * A malformed file is NKPR_ERR_INVALID.  Node ids must fit int32. */
int nkpr_nkbg_read(const char *path, int32_t with_out_lists, nkpr_nkbg *out);
void nkpr_nkbg_free(nkpr_nkbg *g);
/* Replaces NetworkitBinaryWriter::write (networkit/cpp/io/NetworkitBinaryWriter.cpp:19-350):
 * g's out-lists (out_ptr/out_dst/out_weight; for an undirected graph the adjacency, an edge at
 * both endpoints and a self-loop once) and, when directed, its in-lists (colptr/src/weight)
 * in storage order.  chunks as the reference's writer (default 32, at most nodes);
 * weight_format -1 detects it from the weights as the reference's autoDetect does.  Unlike
 * the reference's writer, deleted slots keep their ids (flag bit 0 clear). */
int nkpr_nkbg_write(const char *path, const nkpr_nkbg *g, uint64_t chunks, int32_t weight_format);

/* ---- Approximate personalised PageRank (scd) ------------------------------------------------
 * A handle over a graph's out-lists (Graph::forNeighborsOf order: directed, the out-edges;
 * undirected, the adjacency with an edge at both endpoints and a self-loop once): ptr[z+1],
 * dst[m] (int32), w[m] (FP64, NULL = unit weights), host (device_ptrs = 0) or device pointers;
 * copied to the device with weightedDegree(u, countSelfLoopsTwice = true) per node.  Offsets
 * that do not rise from 0, ids outside [0, z) and negative or NaN weights are NKPR_ERR_INVALID.
 * The handle holds about 72 z + 12 m bytes of device memory. */
typedef struct nkpr_apr nkpr_apr;
typedef struct {
    uint64_t rounds;   /* frontier rounds                                                    */
    uint64_t pushes;   /* node pushes                                                        */
    uint64_t touched;  /* nodes in the result                                                */
    uint64_t edge_pushes; /* out-edges of the pushed nodes (one mass each)                    */
    double ms_total;
} nkpr_apr_stats;

int nkpr_apr_create(uint64_t z, const uint64_t *ptr, const int32_t *dst, const double *w,
                    int32_t device_ptrs, nkpr_apr **out);
int nkpr_apr_destroy(nkpr_apr *a);
/* Replaces ApproximatePageRank::run(seeds) (networkit/cpp/scd/ApproximatePageRank.cpp:18-61):
 * residual pushes from the seeds (a set: repeats ignored, each with residual 1/|seeds|) until
 * every residual is below eps * weightedDegree; alpha is the loop probability.  The pushes run
 * a frontier at a time on the device (the reference runs a FIFO queue on one thread): the
 * identity p + pr_alpha(r) = pr_alpha(seeds) and the stopping rule are the reference's, the
 * values differ from its within that rule.  As in the reference, the share a push sends
 * through a self-loop is dropped, and a node without out-edges is never queued by a
 * neighbour's add (its residual stays) but, as a seed, is pushed while residual is left.  Residuals are summed exactly (64-bit fixed point, 2^-62 units), so a run repeats bit
 * for bit; calls on one handle must not overlap.  *count receives the number of touched nodes
 * (every node that held residual, the reference's result set); ids_out (ascending), p_out and
 * r_out (nullable, host, at least z entries) their ids, scores and final residuals. */
int nkpr_apr_run(nkpr_apr *a, double alpha, double eps, const uint64_t *seeds, uint64_t nseeds,
                 uint64_t *count, int64_t *ids_out, double *p_out, double *r_out,
                 nkpr_apr_stats *stats);

/* Centrality::ranking: ids of the existing nodes ordered by score descending, ties by id
 * ascending; ids_out and scores_out (nullable) receive *count = number of existing nodes
 * entries.  All arrays are host pointers. */
int nkpr_ranking(uint64_t z, const double *scores, const uint8_t *exists, int64_t *ids_out,
                 double *scores_out, uint64_t *count);

/* Message of the last failed call on this thread ("" if none). */
const char *nkpr_last_error(void);

int nkpr_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NKPR_H */
