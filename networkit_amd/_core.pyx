# distutils: language = c++
# cython: language_level=3, boundscheck=False, wraparound=False
"""Cython bindings of the MI355X PageRank drop-in.

Mirrors the reference's Python surface for this path:
  networkit.graph.Graph            (networkit/graph.pyx:29-547, subset)
  networkit.base.Algorithm         (networkit/base.pyx:3-69)
  networkit.centrality.Centrality  (networkit/centrality.pyx:24-100)
  networkit.centrality.PageRank    (networkit/centrality.pyx:1989-2037) and Norm (:1970-1979)
run() releases the GIL around the C++ call, as base.pyx:28-29 does.
"""
from libc.stdint cimport uint64_t, int64_t, int32_t
from libcpp cimport bool as bool_t
from libcpp.vector cimport vector
from libcpp.utility cimport pair
from libcpp.string cimport string
from libcpp.set cimport set as cset
from cython.operator cimport dereference as deref, preincrement as inc

import numpy as np
cimport numpy as cnp

cnp.import_array()

ctypedef uint64_t node
ctypedef uint64_t count
ctypedef double edgeweight

cdef extern from "nkpr.h":
    ctypedef struct nkpr_stats:
        uint64_t iterations
        uint64_t sweeps_fp32
        uint64_t sweeps_fp64
        uint64_t refinements
        uint64_t sweeps_fp32_correction
        double last_step
        double last_step_l1
        double l1_error_bound
        double maximum
        double score_sum
        double ms_sweeps_fp32
        double ms_sweeps_fp64
        double ms_total
    ctypedef struct nkpr_graph_info:
        uint64_t z
        uint64_t n
        uint64_t m
        uint64_t sinks
        uint64_t max_in_degree
        uint64_t work_items
        uint64_t device_bytes
        int32_t weighted
        int32_t device
        double prep_ms

cdef extern from "Graph.hpp" namespace "NetworKit":
    cdef cppclass _Graph "NetworKit::Graph":
        _Graph() except +
        _Graph(count, bool_t, bool_t) except +
        node addNode() except +
        node addNodes(count) except +
        void removeNode(node) except +
        void restoreNode(node) except +
        bool_t hasNode(node)
        void addEdge(node, node, edgeweight) except +
        bool_t hasEdge(node, node)
        void removeEdge(node, node) except +
        edgeweight weight(node, node) except +
        count numberOfNodes()
        count numberOfEdges()
        count upperNodeIdBound()
        count numberOfSelfLoops()
        bool_t isWeighted()
        bool_t isDirected()
        count degree(node) except +
        count degreeIn(node) except +
        count degreeOut(node) except +
        edgeweight weightedDegree(node, bool_t) except +

cdef extern from "PageRank.hpp" namespace "NetworKit":
    cdef cppclass _Algorithm "NetworKit::Algorithm":
        void run() except + nogil
        bool_t hasFinished()
    cdef cppclass _Centrality "NetworKit::Centrality"(_Algorithm):
        vector[double] scores() except +
        vector[pair[node, double]] ranking() except +
        double score(node) except +
        double maximum() except +
        double centralization() except +
    cdef enum _Norm "NetworKit::PageRank::Norm":
        _L1Norm "NetworKit::PageRank::L1Norm"
        _L2Norm "NetworKit::PageRank::L2Norm"
    cdef enum _SinkHandling "NetworKit::PageRank::SinkHandling":
        _NoSinkHandling "NetworKit::PageRank::NoSinkHandling"
        _DistributeSinks "NetworKit::PageRank::DistributeSinks"
    cdef enum _Precision "NetworKit::PageRank::Precision":
        _FP64 "NetworKit::PageRank::FP64"
        _Mixed "NetworKit::PageRank::Mixed"
        _FP32 "NetworKit::PageRank::FP32"
    cdef cppclass _PageRank "NetworKit::PageRank"(_Centrality):
        _PageRank(_Graph, double, double, bool_t, _SinkHandling) except +
        count numberOfIterations() except +
        count maxIterations
        _Norm norm
        _Precision precision
        bool_t certifiedStop
        bool_t timeSweeps
        const nkpr_stats& lastStats()
        const nkpr_graph_info& deviceGraphInfo() except +

cdef extern from "EigenvectorCentrality.hpp" namespace "NetworKit":
    cdef enum _EdgeDirection "NetworKit::EdgeDirection":
        _InEdges "NetworKit::InEdges"
        _OutEdges "NetworKit::OutEdges"
    cdef cppclass _EigenvectorCentrality "NetworKit::EigenvectorCentrality"(_Centrality):
        _EigenvectorCentrality(_Graph, double) except +
        const nkpr_stats& lastStats()
    cdef cppclass _KatzCentrality "NetworKit::KatzCentrality"(_Centrality):
        _KatzCentrality(_Graph, double, double, double) except +
        _EdgeDirection edgeDirection
        const nkpr_stats& lastStats()

cdef extern from "AlgebraicPageRank.hpp" namespace "NetworKit":
    cdef cppclass _AlgebraicPageRank "NetworKit::AlgebraicPageRank<NetworKit::CSRMatrix>"(_Centrality):
        _AlgebraicPageRank(_Graph, double, double) except +
        count numberOfIterations() except +

cdef extern from "NetworkitBinaryIO.hpp" namespace "NetworKit":
    cdef cppclass _NetworkitBinaryReader "NetworKit::NetworkitBinaryReader":
        _NetworkitBinaryReader()
        _Graph read(string) except + nogil
    cdef enum _NetworkitBinaryWeights "NetworKit::NetworkitBinaryWeights":
        pass
    cdef cppclass _NetworkitBinaryWriter "NetworKit::NetworkitBinaryWriter":
        _NetworkitBinaryWriter(uint64_t, _NetworkitBinaryWeights)
        void write(_Graph, string) except + nogil

cdef extern from "SelectiveCommunity.hpp" namespace "NetworKit":
    cdef cppclass _ApproximatePageRank "NetworKit::ApproximatePageRank":
        _ApproximatePageRank(_Graph, double, double) except +
        vector[pair[node, double]] run(cset[node]) except + nogil
    cdef cppclass _PageRankNibble "NetworKit::PageRankNibble":
        _PageRankNibble(_Graph, double, double) except +
        cset[node] expandOneCommunity(cset[node]) except + nogil


cdef class Graph:
    """Graph(n=0, weighted=False, directed=False): NetworKit-compatible graph (subset)."""
    cdef _Graph* _this

    def __cinit__(self, n=0, bool_t weighted=False, bool_t directed=False):
        self._this = new _Graph(n, weighted, directed)

    def __dealloc__(self):
        if self._this != NULL:
            del self._this
            self._this = NULL

    def addNode(self):
        return self._this.addNode()

    def addNodes(self, count k):
        return self._this.addNodes(k)

    def removeNode(self, node u):
        self._this.removeNode(u)

    def restoreNode(self, node u):
        self._this.restoreNode(u)

    def hasNode(self, node u):
        return self._this.hasNode(u)

    def addEdge(self, node u, node v, edgeweight w=1.0):
        self._this.addEdge(u, v, w)
        return self

    def addEdges(self, u, v, w=None):
        """Bulk addEdge(u[i], v[i], w[i]) in order (arrays of equal length)."""
        cdef cnp.ndarray[cnp.uint64_t, ndim=1] uu = np.ascontiguousarray(u, dtype=np.uint64)
        cdef cnp.ndarray[cnp.uint64_t, ndim=1] vv = np.ascontiguousarray(v, dtype=np.uint64)
        cdef cnp.ndarray[cnp.float64_t, ndim=1] ww
        cdef Py_ssize_t i, k = uu.shape[0]
        if vv.shape[0] != k:
            raise ValueError("u and v differ in length")
        if w is None:
            for i in range(k):
                self._this.addEdge(uu[i], vv[i], 1.0)
        else:
            ww = np.ascontiguousarray(w, dtype=np.float64)
            if ww.shape[0] != k:
                raise ValueError("w differs in length")
            for i in range(k):
                self._this.addEdge(uu[i], vv[i], ww[i])
        return self

    def hasEdge(self, node u, node v):
        return self._this.hasEdge(u, v)

    def removeEdge(self, node u, node v):
        self._this.removeEdge(u, v)
        return self

    def weight(self, node u, node v):
        return self._this.weight(u, v)

    def numberOfNodes(self):
        return self._this.numberOfNodes()

    def numberOfEdges(self):
        return self._this.numberOfEdges()

    def upperNodeIdBound(self):
        return self._this.upperNodeIdBound()

    def numberOfSelfLoops(self):
        return self._this.numberOfSelfLoops()

    def isWeighted(self):
        return self._this.isWeighted()

    def isDirected(self):
        return self._this.isDirected()

    def degree(self, node u):
        return self._this.degree(u)

    def degreeIn(self, node u):
        return self._this.degreeIn(u)

    def degreeOut(self, node u):
        return self._this.degreeOut(u)

    def weightedDegree(self, node u, bool_t countSelfLoopsTwice=False):
        return self._this.weightedDegree(u, countSelfLoopsTwice)


cdef class Algorithm:
    """Abstract base class for algorithms (networkit/base.pyx:3-69)."""
    cdef _Algorithm* _this

    def __init__(self, *args, **namedargs):
        if type(self) == Algorithm:
            raise RuntimeError("Error, you may not use Algorithm directly, use a sub-class instead")

    def __cinit__(self, *args, **namedargs):
        self._this = NULL

    def __dealloc__(self):
        if self._this != NULL:
            del self._this
        self._this = NULL

    def run(self):
        if self._this == NULL:
            raise RuntimeError("Error, object not properly initialized")
        with nogil:
            self._this.run()
        return self

    def hasFinished(self):
        if self._this == NULL:
            raise RuntimeError("Error, object not properly initialized")
        return self._this.hasFinished()


cdef class Centrality(Algorithm):
    """Abstract base class for centrality measures (networkit/centrality.pyx:24-100)."""
    cdef Graph _G

    def __init__(self, *args, **kwargs):
        if type(self) == Centrality:
            raise RuntimeError("Error, you may not use Centrality directly, use a sub-class instead")

    def __dealloc__(self):
        self._G = None

    def scores(self):
        if self._this == NULL:
            raise RuntimeError("Error, object not properly initialized")
        return (<_Centrality*>(self._this)).scores()

    def score(self, v):
        if self._this == NULL:
            raise RuntimeError("Error, object not properly initialized")
        return (<_Centrality*>(self._this)).score(v)

    def ranking(self):
        if self._this == NULL:
            raise RuntimeError("Error, object not properly initialized")
        return (<_Centrality*>(self._this)).ranking()

    def maximum(self):
        if self._this == NULL:
            raise RuntimeError("Error, object not properly initialized")
        return (<_Centrality*>(self._this)).maximum()

    def centralization(self):
        if self._this == NULL:
            raise RuntimeError("Error, object not properly initialized")
        return (<_Centrality*>(self._this)).centralization()


class Norm(object):
    """Stopping-criterion norm (PageRank::Norm)."""
    l1norm = 0
    l2norm = 1
    L1Norm = 0
    L2Norm = 1


class SinkHandling(object):
    """What happens to the rank of nodes without out-edges."""
    NoSinkHandling = 0
    DistributeSinks = 1


class Precision(object):
    """Arithmetic of the device sweeps."""
    FP64 = 0
    Mixed = 1
    FP32 = 2


cdef class PageRank(Centrality):
    """PageRank(G, damp=0.85, tol=1e-8, normalized=False, distributeSinks=SinkHandling.NoSinkHandling)

    Compute PageRank as node centrality measure on the MI355X.  Parameters and results follow
    networkit.centrality.PageRank.  Extra properties: `precision` (Precision.FP64 by default,
    Precision.Mixed for FP32 sweeps with FP64 refinement to a certified L1 error <= tol),
    `certifiedStop`, `timeSweeps` and `stats()`.
    """

    def __cinit__(self, Graph G, double damp=0.85, double tol=1e-8, bool_t normalized=False,
                  distributeSinks=SinkHandling.NoSinkHandling):
        self._G = G
        self._this = new _PageRank(G._this[0], damp, tol, normalized,
                                   <_SinkHandling>(<int>distributeSinks))

    def numberOfIterations(self):
        return (<_PageRank*>(self._this)).numberOfIterations()

    property norm:
        def __get__(self):
            return <int>(<_PageRank*>(self._this)).norm
        def __set__(self, norm):
            if int(norm) not in (0, 1):
                raise ValueError("norm must be Norm.l1norm or Norm.l2norm")
            (<_PageRank*>(self._this)).norm = <_Norm>(<int>norm)

    property maxIterations:
        def __get__(self):
            return (<_PageRank*>(self._this)).maxIterations
        def __set__(self, maxIterations):
            if maxIterations < 0:
                raise Exception("Max iterations cannot be a negative number.")
            (<_PageRank*>(self._this)).maxIterations = maxIterations

    property precision:
        def __get__(self):
            return <int>(<_PageRank*>(self._this)).precision
        def __set__(self, p):
            if int(p) not in (0, 1, 2):
                raise ValueError("precision must be Precision.FP64, .Mixed or .FP32")
            (<_PageRank*>(self._this)).precision = <_Precision>(<int>p)

    property certifiedStop:
        def __get__(self):
            return (<_PageRank*>(self._this)).certifiedStop
        def __set__(self, bool_t v):
            (<_PageRank*>(self._this)).certifiedStop = v

    property timeSweeps:
        def __get__(self):
            return (<_PageRank*>(self._this)).timeSweeps
        def __set__(self, bool_t v):
            (<_PageRank*>(self._this)).timeSweeps = v

    def stats(self):
        """Statistics of the last run (sweeps per precision, error bound, timings)."""
        cdef nkpr_stats s = (<_PageRank*>(self._this)).lastStats()
        return dict(iterations=s.iterations, sweeps_fp32=s.sweeps_fp32,
                    sweeps_fp64=s.sweeps_fp64, refinements=s.refinements,
                    sweeps_fp32_correction=s.sweeps_fp32_correction,
                    last_step=s.last_step, last_step_l1=s.last_step_l1,
                    l1_error_bound=s.l1_error_bound, maximum=s.maximum,
                    score_sum=s.score_sum, ms_sweeps_fp32=s.ms_sweeps_fp32,
                    ms_sweeps_fp64=s.ms_sweeps_fp64, ms_total=s.ms_total)

    def deviceGraphInfo(self):
        cdef nkpr_graph_info i = (<_PageRank*>(self._this)).deviceGraphInfo()
        return dict(z=i.z, n=i.n, m=i.m, sinks=i.sinks, max_in_degree=i.max_in_degree,
                    work_items=i.work_items, device_bytes=i.device_bytes,
                    weighted=bool(i.weighted), device=i.device, prep_ms=i.prep_ms)


cdef dict _power_stats(nkpr_stats s):
    return dict(iterations=s.iterations, last_step=s.last_step, norm=s.score_sum,
                ms_total=s.ms_total)


cdef class EigenvectorCentrality(Centrality):
    """EigenvectorCentrality(G, tol=1e-9)

    The leading eigenvector of the graph's adjacency matrix, normalised in the 2-norm
    (networkit.centrality.EigenvectorCentrality, centrality.pyx:1950-1968), computed on the
    MI355X.  `tol` bounds the change of the norm between iterations."""

    def __cinit__(self, Graph G, double tol=1e-9):
        self._G = G
        self._this = new _EigenvectorCentrality(G._this[0], tol)

    def stats(self):
        """Statistics of the last run: iterations, final change and value of the norm."""
        return _power_stats((<_EigenvectorCentrality*>(self._this)).lastStats())


class EdgeDirection(object):
    """Which neighbours Katz centrality sums over (KatzCentrality.hpp:17)."""
    inEdges = 0
    outEdges = 1


cdef class KatzCentrality(Centrality):
    """KatzCentrality(G, alpha=0, beta=0.1, tol=1e-8)

    Katz centrality (networkit.centrality.KatzCentrality, centrality.pyx:1678-1710) computed
    on the MI355X.  alpha = 0 takes 1 / (1 + the largest out-degree); `edgeDirection` selects
    in-neighbours (default) or out-neighbours."""

    def __cinit__(self, Graph G, alpha=0, beta=0.1, tol=1e-8):
        self._G = G
        self._this = new _KatzCentrality(G._this[0], alpha, beta, tol)

    property edgeDirection:
        def __get__(self):
            return <int>(<_KatzCentrality*>(self._this)).edgeDirection
        def __set__(self, edgeDirection):
            if int(edgeDirection) not in (0, 1):
                raise ValueError("edgeDirection must be EdgeDirection.inEdges or .outEdges")
            (<_KatzCentrality*>(self._this)).edgeDirection = <_EdgeDirection>(<int>edgeDirection)

    def stats(self):
        """Statistics of the last run: iterations, final change and value of the norm."""
        return _power_stats((<_KatzCentrality*>(self._this)).lastStats())


cdef class AlgebraicPageRank(Centrality):
    """AlgebraicPageRank(G, damp=0.85, tol=1e-8)

    PageRank on the algebraic module's CSR matrix (C++ NetworKit::AlgebraicPageRank<CSRMatrix>,
    AlgebraicPageRank.hpp:25-150; not exported by NetworKit's Python package): the matrix
    M = damp A^T D^-1 is built on the host at construction, run() iterates M x + (1-damp)/n on
    the MI355X until the L2 step is at most tol, then divides by the sum.  Every slot of the
    graph is a row (deleted nodes included, as in the reference); maximum() is 1."""

    def __cinit__(self, Graph G, double damp=0.85, double tol=1e-8):
        self._G = G
        self._this = new _AlgebraicPageRank(G._this[0], damp, tol)

    def numberOfIterations(self):
        return (<_AlgebraicPageRank*>(self._this)).numberOfIterations()


def _read_nkbg(path):
    """NetworkitBinaryReader::read through the C++ drop-in: a Graph from an nkbg002 file."""
    cdef _NetworkitBinaryReader reader
    cdef string p = path.encode()
    cdef Graph G = Graph(0)
    with nogil:
        G._this[0] = reader.read(p)
    return G


def _write_nkbg(Graph G, path, uint64_t chunks=32, int weightsType=5):
    """NetworkitBinaryWriter(chunks, weightsType).write(G, path) through the C++ drop-in."""
    cdef _NetworkitBinaryWriter* w = new _NetworkitBinaryWriter(chunks, <_NetworkitBinaryWeights>weightsType)
    cdef string p = path.encode()
    try:
        with nogil:
            w.write(G._this[0], p)
    finally:
        del w


cdef cset[node] _node_set(seeds):
    cdef cset[node] out
    cdef node x
    if isinstance(seeds, int):
        seeds = [seeds]
    for s in seeds:
        x = s
        out.insert(x)
    return out


cdef class ApproximatePageRank:
    """ApproximatePageRank(G, alpha, epsilon=1e-12)

    Approximate personalised PageRank from seed nodes by residual pushes
    (networkit.scd.ApproximatePageRank, scd.pyx:93-141; ApproximatePageRank.cpp:18-61), the
    pushes running a frontier at a time on the MI355X.  run(seeds) returns the [(node, score)]
    pairs of every node the pushes touched (ids ascending)."""
    cdef _ApproximatePageRank* _this
    cdef Graph _G

    def __cinit__(self, Graph G not None, double alpha, double epsilon=1e-12):
        self._G = G
        self._this = new _ApproximatePageRank(G._this[0], alpha, epsilon)

    def __dealloc__(self):
        if self._this != NULL:
            del self._this
        self._this = NULL

    def run(self, seeds):
        cdef cset[node] sd = _node_set(seeds)
        cdef vector[pair[node, double]] res
        with nogil:
            res = self._this.run(sd)
        return [(p.first, p.second) for p in res]


cdef class PageRankNibble:
    """PageRankNibble(G, alpha, epsilon)

    A community around seed nodes: the approximate personalised PageRank (pushes on the MI355X)
    and its best sweep set by conductance (networkit.scd.PageRankNibble, scd.pyx:143-156;
    PageRankNibble.cpp:28-96).  run(seeds) -> {seed: set of nodes}, one community per seed
    (SelectiveCommunityDetector.cpp:15-23); expandOneCommunity(seeds) -> set of nodes."""
    cdef _PageRankNibble* _this
    cdef Graph _G

    def __cinit__(self, Graph G not None, double alpha, double epsilon):
        self._G = G
        self._this = new _PageRankNibble(G._this[0], alpha, epsilon)

    def __dealloc__(self):
        if self._this != NULL:
            del self._this
        self._this = NULL

    def expandOneCommunity(self, seeds):
        cdef cset[node] sd = _node_set(seeds)
        cdef cset[node] res
        with nogil:
            res = self._this.expandOneCommunity(sd)
        cdef list out = []
        cdef cset[node].iterator it = res.begin()
        while it != res.end():
            out.append(deref(it))
            inc(it)
        return set(out)

    def run(self, seeds):
        if isinstance(seeds, int):
            seeds = [seeds]
        return {s: self.expandOneCommunity([s]) for s in sorted(set(seeds))}
