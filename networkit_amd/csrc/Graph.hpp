// NetworKit-compatible Graph (the subset PageRank needs), written for this library.
//
// Storage follows NetworKit's semantics (include/networkit/graph/Graph.hpp:97-146): per-node
// adjacency vectors, weights in the same order, an exists[] flag per slot.  Directed graphs
// keep out- and in-lists; undirected graphs keep one list per node in which an edge appears
// at both endpoints and a self-loop once (networkit/cpp/graph/Graph.cpp:549-597).  Multi-
// edges are kept.  exportInLists() flattens the lists PageRank pulls over
// (Graph::forInEdgesOf, Graph.hpp:1918-1938) into the CSC the device library consumes.
#pragma once

#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace NetworKit {

using node = uint64_t;
using index = uint64_t;
using count = uint64_t;
using edgeweight = double;
constexpr edgeweight defaultEdgeWeight = 1.0;
constexpr node none = std::numeric_limits<node>::max();

class Graph {
public:
    explicit Graph(count n = 0, bool weighted = false, bool directed = false);

    node addNode();
    node addNodes(count k);                    // returns the last new id
    void removeNode(node v);                   // removes incident edges first (Graph.cpp:514-527)
    void restoreNode(node v);
    bool hasNode(node v) const noexcept { return v < z && exists[v]; }

    void addEdge(node u, node v, edgeweight ew = defaultEdgeWeight);
    bool hasEdge(node u, node v) const noexcept;
    void removeEdge(node u, node v);           // removes one occurrence
    edgeweight weight(node u, node v) const;

    count numberOfNodes() const noexcept { return n; }
    count numberOfEdges() const noexcept { return m; }
    count upperNodeIdBound() const noexcept { return z; }
    count numberOfSelfLoops() const noexcept { return selfLoops; }
    bool isWeighted() const noexcept { return weighted; }
    bool isDirected() const noexcept { return directed; }

    // The accessors taking a node id throw std::out_of_range for an id >= upperNodeIdBound()
    // (NetworKit leaves that undefined; the Python layer turns it into IndexError).
    count degree(node v) const { return outEdges[checked(v)].size(); }
    count degreeOut(node v) const { return outEdges[checked(v)].size(); }
    count degreeIn(node v) const {
        return directed ? inEdges[checked(v)].size() : outEdges[checked(v)].size();
    }
    // Graph::computeWeightedDegree(u, inDegree = false, countSelfLoopsTwice), Graph.cpp:427-455
    edgeweight weightedDegree(node u, bool countSelfLoopsTwice = false) const;

    template <typename L> void forNodes(L handle) const {
        for (node v = 0; v < z; ++v)
            if (exists[v]) handle(v);
    }

    // Graph::forNeighborsOf (Graph.hpp:1906-1915 with every edge used): u's out-list (the
    // adjacency when undirected) in storage order.  handle(v, weight).
    template <typename L> void forNeighborsOf(node u, L handle) const {
        const auto &lst = outEdges[checked(u)];
        for (size_t i = 0; i < lst.size(); ++i)
            handle(lst[i], weighted ? outEdgeWeights[u][i] : defaultEdgeWeight);
    }

    // Graph::totalEdgeWeight (Graph.cpp:860-868): the sum over forEdges, or m unweighted.
    edgeweight totalEdgeWeight() const noexcept {
        if (!weighted) return static_cast<edgeweight>(m) * defaultEdgeWeight;
        edgeweight sum = 0.0;
        forEdges([&](node, node, edgeweight w) { sum += w; });
        return sum;
    }

    // Graph::forEdges (Graph.hpp:1941-1945, 1895-1915): u ascending, each out-list in
    // storage order; an undirected edge is visited once, at its endpoint u >= v.
    // handle(u, v, weight).
    template <typename L> void forEdges(L handle) const {
        for (node u = 0; u < z; ++u) {
            const auto &lst = outEdges[u];
            for (size_t i = 0; i < lst.size(); ++i) {
                const node v = lst[i];
                if (directed || u >= v)
                    handle(u, v, weighted ? outEdgeWeights[u][i] : defaultEdgeWeight);
            }
        }
    }

    // Flattened pull lists: row u holds the in-neighbours of u (directed) or the adjacency of
    // u (undirected), in insertion order.  Weights are filled only for weighted graphs.
    void exportInLists(std::vector<uint64_t> &colptr, std::vector<int32_t> &src,
                       std::vector<float> *w32, std::vector<double> *w64) const {
        exportLists(false, colptr, src, w32, w64);
    }
    // outgoing = true: row u holds u's out-neighbours instead (the same lists when undirected)
    void exportLists(bool outgoing, std::vector<uint64_t> &colptr, std::vector<int32_t> &src,
                     std::vector<float> *w32, std::vector<double> *w64) const;
    const std::vector<bool> &existsFlags() const noexcept { return exists; }

    // A graph built from flat lists in storage order (what a file reader decodes): for an
    // undirected graph `pull` holds each node's adjacency; for a directed one `out` holds the
    // out-lists and `pull` the in-lists.  Weights may be NULL (unweighted); exists may be NULL
    // (every slot is a node).  edges and selfLoops are the counts the lists describe.
    static Graph fromLists(count z, bool weighted, bool directed, const uint8_t *exists,
                           const uint64_t *outPtr, const int32_t *outIds, const double *outW,
                           const uint64_t *pullPtr, const int32_t *pullIds, const double *pullW,
                           count edges, count selfLoops);
    // The out-lists in storage order (for an undirected graph, the adjacency).
    void exportOutLists(std::vector<uint64_t> &ptr, std::vector<int32_t> &ids,
                        std::vector<double> *w) const {
        exportLists(true, ptr, ids, nullptr, w);
    }

    // Increments on every structural or weight change; consumers cache device copies by it.
    uint64_t version() const noexcept { return ver; }

private:
    node checked(node v) const {
        if (v >= z) throw std::out_of_range("node id " + std::to_string(v) + " >= " +
                                            std::to_string(z));
        return v;
    }
    count n, m, z, selfLoops;
    bool weighted, directed;
    uint64_t ver = 0;
    std::vector<bool> exists;
    std::vector<std::vector<node>> outEdges, inEdges;
    std::vector<std::vector<edgeweight>> outEdgeWeights, inEdgeWeights;
};

}  // namespace NetworKit
