"""CPU oracle for NetworKit's PageRank -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only as the checker (or, for the baseline, as the thing timed on the
host cores).  The product (``networkit_amd``) never imports it and has no CPU fallback.

The arithmetic lives in ``pagerank_oracle.c`` (a restatement of
networkit/cpp/centrality/PageRank.cpp:20-104), ``power_oracle.c`` (EigenvectorCentrality.cpp:
21-66, KatzCentrality.cpp:28-76) and ``algebraic_oracle.c`` (AlgebraicPageRank.hpp:35-119,
CSRGeneralMatrix.hpp:151-182,563-578) and ``scd_oracle.c`` (ApproximatePageRank.cpp:18-61,
PageRankNibble.cpp:28-89); this module loads it with ctypes and adds the
NetworKit-semantics edge-list -> in-list conversion (Graph.cpp:549-597, Graph.hpp:1918-1938)
that the C code consumes.

Parity status of each function is pinned in tests/test_oracle.py:
  pagerank        reference known answers (CentralityGTest.cpp:484-664), sparse direct solve,
                  dense power iteration, invariants; with a teleport vector: sparse direct
                  solve with that right-hand side, uniform teleport == the reference loop,
                  linearity in the teleport vector
  ranking         brute-force sort
  eigenvector     8-node known answers (CentralityGTest.cpp:541-578), dense eigendecomposition
                  of the adjacency matrix (symmetric and directed), regular-graph invariant
  katz            wiki-Vote top-1 with alpha 5e-4 (CentralityGTest.cpp:216-225), sparse direct
                  solve of (I - alpha A) x = alpha A 1 + beta 1, regular-graph invariant
  in_lists        hand-built NetworKit adjacency of small graphs (directed, undirected,
                  self-loops, multi-edges)
  for_edges       hand-built forEdges order of small graphs (directed, undirected, self-loops)
  csr_spmv        scipy.sparse products
  algebraic_matrix  dense construction of d A^T D^-1 from the edge list
  algebraic_pagerank  reference known answers (AlgebraicPageRankGTest.cpp:21-94), sparse direct
                  solve of (I - M) x = (1-d)/n 1, equality with pagerank() on graphs without
                  deleted nodes
  approximate_pagerank  the ACL identity p + pr(r) = pr(seeds) against a sparse direct solve
                  of the lazy personalised PageRank, the stopping invariant, mass conservation
  best_sweep_set  brute force over every prefix with set-based cut and volume
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_pagerank.so")
_lib = None

L1 = 0
L2 = 1


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    p = ctypes.c_void_p
    u64 = ctypes.c_uint64
    lib.oracle_pagerank.argtypes = [u64, p, p, p, p, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, u64, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, p, ctypes.POINTER(u64),
                                    ctypes.POINTER(ctypes.c_double)]
    lib.oracle_pagerank.restype = ctypes.c_int
    lib.oracle_pagerank_tv.argtypes = [u64, p, p, p, p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, u64, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, p, p, ctypes.POINTER(u64),
                                       ctypes.POINTER(ctypes.c_double)]
    lib.oracle_pagerank_tv.restype = ctypes.c_int
    lib.oracle_ranking.argtypes = [u64, p, p, p]
    lib.oracle_ranking.restype = u64
    lib.oracle_out_degrees.argtypes = [u64, p, p, p, p, p]
    lib.oracle_out_degrees.restype = None
    lib.oracle_eigenvector.argtypes = [u64, p, p, p, p, ctypes.c_double, u64, p,
                                       ctypes.POINTER(u64)]
    lib.oracle_eigenvector.restype = ctypes.c_int
    lib.oracle_katz.argtypes = [u64, p, p, p, p, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, u64, p, ctypes.POINTER(u64)]
    lib.oracle_katz.restype = ctypes.c_int
    lib.oracle_katz_default_alpha.argtypes = [u64, p, p, ctypes.c_int]
    lib.oracle_katz_default_alpha.restype = ctypes.c_double
    lib.oracle_csr_spmv.argtypes = [u64, p, p, p, p, p]
    lib.oracle_csr_spmv.restype = None
    lib.oracle_algebraic_nnz.argtypes = [u64, p, p, ctypes.c_int]
    lib.oracle_algebraic_nnz.restype = u64
    lib.oracle_algebraic_matrix.argtypes = [u64, u64, p, p, p, ctypes.c_int, ctypes.c_double,
                                            p, p, p]
    lib.oracle_algebraic_matrix.restype = ctypes.c_int
    lib.oracle_algebraic_pagerank.argtypes = [u64, p, p, p, ctypes.c_double, ctypes.c_double,
                                              u64, p, ctypes.POINTER(u64)]
    lib.oracle_algebraic_pagerank.restype = ctypes.c_int
    lib.oracle_apr.argtypes = [u64, p, p, p, ctypes.c_double, ctypes.c_double, p, u64, p, p, p]
    lib.oracle_apr.restype = ctypes.c_int64
    lib.oracle_sweep.argtypes = [u64, p, p, p, u64, ctypes.c_double, u64, p, p, p,
                                 ctypes.POINTER(ctypes.c_double)]
    lib.oracle_sweep.restype = u64
    lib.oracle_num_threads.argtypes = []
    lib.oracle_num_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class InLists:
    """Flattened NetworKit in-lists: in_ptr[z+1], in_src[m] (uint64), in_w[m] (float64 or
    None), exists[z] (uint8 or None)."""

    def __init__(self, z, in_ptr, in_src, in_w=None, exists=None, directed=True):
        self.z = int(z)
        self.in_ptr = np.ascontiguousarray(in_ptr, dtype=np.uint64)
        self.in_src = np.ascontiguousarray(in_src, dtype=np.uint64)
        self.in_w = None if in_w is None else np.ascontiguousarray(in_w, dtype=np.float64)
        self.exists = None if exists is None else np.ascontiguousarray(exists, dtype=np.uint8)
        self.directed = directed

    @property
    def m(self):
        return int(self.in_ptr[-1])


def in_lists(z, u, v, w=None, directed=True, exists=None):
    """Build the in-lists NetworKit's Graph would hold after ``addEdge(u[i], v[i], w[i])`` in
    order i = 0, 1, ...  (Graph.cpp:549-597).  Directed: inEdges[v] gets u.  Undirected:
    outEdges[u] gets v and, unless u == v, outEdges[v] gets u (a self-loop is stored once).
    The relative order of entries within a list is insertion order, as in NetworKit."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    wt = None if w is None else np.asarray(w, dtype=np.float64)
    if directed:
        rows, srcs, seq = v, u, np.arange(len(u), dtype=np.int64)
        ws = wt
    else:
        loop = u == v
        k = len(u)
        # event order: edge i contributes (row u, src v) then (row v, src u) if not a loop
        rows = np.concatenate([u, v[~loop]])
        srcs = np.concatenate([v, u[~loop]])
        seq = np.concatenate([2 * np.arange(k), 2 * np.arange(k)[~loop] + 1])
        ws = None if wt is None else np.concatenate([wt, wt[~loop]])
    order = np.lexsort((seq, rows))
    rows_s = rows[order]
    counts = np.bincount(rows_s, minlength=z) if len(rows_s) else np.zeros(z, np.int64)
    ptr = np.zeros(z + 1, dtype=np.uint64)
    ptr[1:] = np.cumsum(counts)
    src = srcs[order].astype(np.uint64)
    wout = None if ws is None else ws[order]
    return InLists(z, ptr, src, wout, exists, directed)


def pagerank(g, damp=0.85, tol=1e-8, norm=L2, max_iterations=2**64 - 1, normalized=False,
             sink_mode=0, start=None, threads=0, teleport=None):
    """Run the restated PageRank::run.  ``start`` is the start vector (default 1/n in every
    slot, as a freshly constructed reference object).  ``teleport`` (personalised PageRank, not
    in the reference) replaces (1-d)/n by (1-d)*teleport[u].  Returns (scores, iterations,
    maximum)."""
    lib = _load()
    z = g.z
    n = z if g.exists is None else int(g.exists.sum())
    if start is None:
        scores = np.full(z, 1.0 / max(n, 1), dtype=np.float64)
    else:
        scores = np.array(start, dtype=np.float64, copy=True)
    it = ctypes.c_uint64(0)
    mx = ctypes.c_double(0.0)
    tv = None if teleport is None else np.ascontiguousarray(teleport, dtype=np.float64)
    if tv is not None and tv.shape != (z,):
        raise ValueError("teleport must have one entry per slot")
    rc = lib.oracle_pagerank_tv(z, _ptr(g.exists), _ptr(g.in_ptr), _ptr(g.in_src),
                                _ptr(g.in_w), damp, tol, norm, int(max_iterations),
                                int(bool(normalized)), int(sink_mode), int(threads), _ptr(tv),
                                _ptr(scores), ctypes.byref(it), ctypes.byref(mx))
    if rc != 0:
        raise RuntimeError("oracle_pagerank failed")
    return scores, int(it.value), float(mx.value)


def ranking(scores, exists=None):
    """Existing node ids in Centrality::ranking order (score desc, id asc)."""
    lib = _load()
    scores = np.ascontiguousarray(scores, dtype=np.float64)
    z = len(scores)
    ids = np.zeros(max(z, 1), dtype=np.uint64)
    ex = None if exists is None else np.ascontiguousarray(exists, dtype=np.uint8)
    k = lib.oracle_ranking(z, _ptr(ex), _ptr(scores), _ptr(ids))
    return ids[:k].astype(np.int64)


def out_degrees(g):
    """(weighted out-degree float64[z], unweighted out-count uint64[z])."""
    lib = _load()
    deg = np.zeros(g.z, dtype=np.float64)
    cnt = np.zeros(g.z, dtype=np.uint64)
    lib.oracle_out_degrees(g.z, _ptr(g.in_ptr), _ptr(g.in_src), _ptr(g.in_w), _ptr(deg),
                           _ptr(cnt))
    return deg, cnt


def eigenvector(g, tol=1e-8, max_iterations=2**64 - 1):
    """The restated EigenvectorCentrality::run on in-lists ``g``.  Returns (scores,
    iterations)."""
    lib = _load()
    scores = np.zeros(g.z, dtype=np.float64)
    it = ctypes.c_uint64(0)
    rc = lib.oracle_eigenvector(g.z, _ptr(g.exists), _ptr(g.in_ptr), _ptr(g.in_src),
                                _ptr(g.in_w), tol, int(max_iterations), _ptr(scores),
                                ctypes.byref(it))
    if rc != 0:
        raise RuntimeError("oracle_eigenvector failed")
    return scores, int(it.value)


def katz_default_alpha(g):
    """1 / (1 + largest unweighted out-degree), the reference's alpha = 0 default."""
    return float(_load().oracle_katz_default_alpha(g.z, _ptr(g.in_ptr), _ptr(g.in_src),
                                                   int(bool(g.directed))))


def katz(g, alpha=0.0, beta=0.1, tol=1e-8, max_iterations=2**64 - 1):
    """The restated KatzCentrality::run (in-neighbours; pass the reversed graph's in-lists for
    the OutEdges direction).  alpha = 0 takes the reference default.  Returns (scores,
    iterations)."""
    lib = _load()
    if alpha == 0:
        alpha = katz_default_alpha(g)
    scores = np.zeros(g.z, dtype=np.float64)
    it = ctypes.c_uint64(0)
    rc = lib.oracle_katz(g.z, _ptr(g.exists), _ptr(g.in_ptr), _ptr(g.in_src), _ptr(g.in_w),
                         alpha, beta, tol, int(max_iterations), _ptr(scores), ctypes.byref(it))
    if rc != 0:
        raise RuntimeError("oracle_katz failed")
    return scores, int(it.value)


def for_edges(z, u, v, w=None, directed=True):
    """The edges of the NetworKit Graph built by ``addEdge(u[i], v[i], w[i])`` in order, in
    Graph::forEdges order (Graph.hpp:1941-1945, 1895-1915): u ascending, each out-list in
    insertion order; an undirected edge is visited once, at its endpoint u >= v (the out-list of
    an undirected graph holds an edge at both endpoints, a self-loop once).  Returns (eu, ev,
    ew) with ew None when unweighted."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    wt = None if w is None else np.asarray(w, dtype=np.float64)
    if directed:
        rows, cols, seq, ws = u, v, np.arange(len(u), dtype=np.int64), wt
    else:
        loop = u == v
        k = len(u)
        rows = np.concatenate([u, v[~loop]])
        cols = np.concatenate([v, u[~loop]])
        seq = np.concatenate([2 * np.arange(k), 2 * np.arange(k)[~loop] + 1])
        ws = None if wt is None else np.concatenate([wt, wt[~loop]])
        keep = rows >= cols
        rows, cols, seq = rows[keep], cols[keep], seq[keep]
        ws = None if ws is None else ws[keep]
    order = np.lexsort((seq, rows))
    return (rows[order].astype(np.uint64), cols[order].astype(np.uint64),
            None if ws is None else ws[order])


class CSR:
    """A CSRGeneralMatrix<double> in NetworKit's element types: row_idx[rows+1], col_idx[nnz]
    (uint64), values[nnz] (float64)."""

    def __init__(self, rows, cols, row_idx, col_idx, values):
        self.rows = int(rows)
        self.cols = int(cols)
        self.row_idx = np.ascontiguousarray(row_idx, dtype=np.uint64)
        self.col_idx = np.ascontiguousarray(col_idx, dtype=np.uint64)
        self.values = np.ascontiguousarray(values, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.row_idx[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values, self.col_idx.astype(np.int64),
                              self.row_idx.astype(np.int64)), shape=(self.rows, self.cols))


def csr_spmv(a, x):
    """CSRGeneralMatrix::operator*(const Vector&) (CSRGeneralMatrix.hpp:563-578)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    if x.shape != (a.cols,):
        raise ValueError("x must have one entry per column")
    y = np.zeros(a.rows, dtype=np.float64)
    _load().oracle_csr_spmv(a.rows, _ptr(a.row_idx), _ptr(a.col_idx), _ptr(a.values), _ptr(x),
                            _ptr(y))
    return y


def algebraic_matrix(z, eu, ev, ew=None, directed=True, damp=0.85):
    """The matrix M that AlgebraicPageRank(G, damp) builds (AlgebraicPageRank.hpp:35-54) from
    G's edges in forEdges order (see for_edges)."""
    lib = _load()
    eu = np.ascontiguousarray(eu, dtype=np.uint64)
    ev = np.ascontiguousarray(ev, dtype=np.uint64)
    ew = None if ew is None else np.ascontiguousarray(ew, dtype=np.float64)
    nnz = int(lib.oracle_algebraic_nnz(len(eu), _ptr(eu), _ptr(ev), int(bool(directed))))
    row = np.zeros(z + 1, dtype=np.uint64)
    col = np.zeros(max(nnz, 1), dtype=np.uint64)
    val = np.zeros(max(nnz, 1), dtype=np.float64)
    rc = lib.oracle_algebraic_matrix(z, len(eu), _ptr(eu), _ptr(ev), _ptr(ew),
                                     int(bool(directed)), damp, _ptr(row), _ptr(col), _ptr(val))
    if rc != 0:
        raise RuntimeError("oracle_algebraic_matrix failed")
    return CSR(z, z, row, col[:nnz], val[:nnz])


def algebraic_pagerank(mat, damp=0.85, tol=1e-8, max_iterations=2**64 - 1):
    """The restated AlgebraicPageRank::run (AlgebraicPageRank.hpp:94-119) on M.  Returns
    (scores, iterations)."""
    lib = _load()
    scores = np.zeros(mat.rows, dtype=np.float64)
    it = ctypes.c_uint64(0)
    rc = lib.oracle_algebraic_pagerank(mat.rows, _ptr(mat.row_idx), _ptr(mat.col_idx),
                                       _ptr(mat.values), damp, tol, int(max_iterations),
                                       _ptr(scores), ctypes.byref(it))
    if rc != 0:
        raise RuntimeError("oracle_algebraic_pagerank failed")
    return scores, int(it.value)


def out_lists(z, u, v, w=None, directed=True):
    """NetworKit's out-lists (Graph::forNeighborsOf order) of the Graph built by
    ``addEdge(u[i], v[i], w[i])`` in order: directed, u's out-edges; undirected, u's
    adjacency (an edge at both endpoints, a self-loop once), insertion order.  Returns
    (ptr, dst, w)."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    wt = None if w is None else np.asarray(w, dtype=np.float64)
    if directed:
        rows, cols, seq, ws = u, v, np.arange(len(u), dtype=np.int64), wt
    else:
        loop = u == v
        k = len(u)
        rows = np.concatenate([u, v[~loop]])
        cols = np.concatenate([v, u[~loop]])
        seq = np.concatenate([2 * np.arange(k), 2 * np.arange(k)[~loop] + 1])
        ws = None if wt is None else np.concatenate([wt, wt[~loop]])
    order = np.lexsort((seq, rows))
    ptr = np.zeros(z + 1, dtype=np.uint64)
    ptr[1:] = np.cumsum(np.bincount(rows, minlength=z))
    return (ptr, np.ascontiguousarray(cols[order], dtype=np.uint64),
            None if ws is None else np.ascontiguousarray(ws[order]))


def approximate_pagerank(z, ptr, dst, w, seeds, alpha, eps=1e-12):
    """The restated ApproximatePageRank::run (ApproximatePageRank.cpp:18-61) on out-lists.
    Returns (p, r, touched) as dense arrays of z entries."""
    lib = _load()
    seeds = np.unique(np.asarray(list(seeds), dtype=np.uint64))
    p = np.zeros(z, dtype=np.float64)
    r = np.zeros(z, dtype=np.float64)
    touched = np.zeros(z, dtype=np.uint8)
    ptr = np.ascontiguousarray(ptr, dtype=np.uint64)
    dst = np.ascontiguousarray(dst, dtype=np.uint64)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    k = lib.oracle_apr(z, _ptr(ptr), _ptr(dst), _ptr(w), alpha, eps, _ptr(seeds), len(seeds),
                       _ptr(p), _ptr(r), _ptr(touched))
    if k < 0:
        raise RuntimeError("oracle_apr failed")
    return p, r, touched


def weighted_degree2(z, ptr, dst, w):
    """Graph::weightedDegree(u, countSelfLoopsTwice = true) from out-lists."""
    ptr = np.asarray(ptr, dtype=np.int64)
    rows = np.repeat(np.arange(z), np.diff(ptr))
    wt = np.ones(len(dst)) if w is None else np.asarray(w, dtype=np.float64)
    wt = np.where(np.asarray(dst, dtype=np.int64) == rows, 2.0 * wt, wt)
    return np.bincount(rows, weights=wt, minlength=z)


def best_sweep_set(z, ptr, dst, w, n, total_volume, ids, score):
    """PageRankNibble::bestSweepSet (PageRankNibble.cpp:28-89) on nodes ``ids`` with scores
    ``score`` (p / weightedDegree(u, true)).  Returns (order, best prefix length, best
    conductance)."""
    lib = _load()
    ids = np.ascontiguousarray(ids, dtype=np.uint64)
    score = np.ascontiguousarray(score, dtype=np.float64)
    order = np.zeros(max(len(ids), 1), dtype=np.uint64)
    best = ctypes.c_double(0.0)
    ptr = np.ascontiguousarray(ptr, dtype=np.uint64)
    dst = np.ascontiguousarray(dst, dtype=np.uint64)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    k = lib.oracle_sweep(z, _ptr(ptr), _ptr(dst), _ptr(w), int(n), float(total_volume),
                         len(ids), _ptr(ids), _ptr(score), _ptr(order), ctypes.byref(best))
    return order[:len(ids)].astype(np.int64), int(k), float(best.value)


def num_threads():
    return int(_load().oracle_num_threads())
