#include "Graph.hpp"

#include <algorithm>
#include <string>

namespace NetworKit {

Graph::Graph(count n_, bool weighted_, bool directed_)
    : n(n_), m(0), z(n_), selfLoops(0), weighted(weighted_), directed(directed_),
      exists(n_, true), outEdges(n_), inEdges(directed_ ? n_ : 0),
      outEdgeWeights(weighted_ ? n_ : 0), inEdgeWeights(weighted_ && directed_ ? n_ : 0) {}

node Graph::addNode() { return addNodes(1); }

node Graph::addNodes(count k) {
    z += k;
    n += k;
    exists.resize(z, true);
    outEdges.resize(z);
    if (directed) inEdges.resize(z);
    if (weighted) {
        outEdgeWeights.resize(z);
        if (directed) inEdgeWeights.resize(z);
    }
    ++ver;
    return z - 1;
}

void Graph::removeNode(node v) {
    if (!hasNode(v)) throw std::runtime_error("removeNode: no such node");
    while (!outEdges[v].empty()) removeEdge(v, outEdges[v].front());
    if (directed)
        while (!inEdges[v].empty()) removeEdge(inEdges[v].front(), v);
    exists[v] = false;
    --n;
    ++ver;
}

void Graph::restoreNode(node v) {
    if (v >= z || exists[v]) throw std::runtime_error("restoreNode: slot is not a deleted node");
    exists[v] = true;
    ++n;
    ++ver;
}

void Graph::addEdge(node u, node v, edgeweight ew) {
    if (!hasNode(u) || !hasNode(v)) throw std::runtime_error("addEdge: endpoint is not a node");
    ++m;
    outEdges[u].push_back(v);
    if (directed) {
        inEdges[v].push_back(u);
        if (weighted) {
            inEdgeWeights[v].push_back(ew);
            outEdgeWeights[u].push_back(ew);
        }
    } else if (u == v) {
        if (weighted) outEdgeWeights[u].push_back(ew);
    } else {
        outEdges[v].push_back(u);
        if (weighted) {
            outEdgeWeights[u].push_back(ew);
            outEdgeWeights[v].push_back(ew);
        }
    }
    if (u == v) ++selfLoops;
    ++ver;
}

bool Graph::hasEdge(node u, node v) const noexcept {
    if (u >= z || v >= z) return false;
    if (!directed && outEdges[u].size() > outEdges[v].size())
        return std::find(outEdges[v].begin(), outEdges[v].end(), u) != outEdges[v].end();
    if (directed && outEdges[u].size() > inEdges[v].size())
        return std::find(inEdges[v].begin(), inEdges[v].end(), u) != inEdges[v].end();
    return std::find(outEdges[u].begin(), outEdges[u].end(), v) != outEdges[u].end();
}

namespace {
template <typename T> void erase_at(std::vector<T> &vec, size_t i) {
    vec[i] = vec.back();
    vec.pop_back();
}
size_t find_pos(const std::vector<node> &lst, node x) {
    for (size_t i = 0; i < lst.size(); ++i)
        if (lst[i] == x) return i;
    return lst.size();
}
}  // namespace

void Graph::removeEdge(node u, node v) {
    if (u >= z || v >= z) throw std::runtime_error("removeEdge: no such edge");
    const size_t i = find_pos(outEdges[u], v);
    if (i == outEdges[u].size()) throw std::runtime_error("removeEdge: no such edge");
    erase_at(outEdges[u], i);
    if (weighted) erase_at(outEdgeWeights[u], i);
    if (directed) {
        const size_t j = find_pos(inEdges[v], u);
        erase_at(inEdges[v], j);
        if (weighted) erase_at(inEdgeWeights[v], j);
    } else if (u != v) {
        const size_t j = find_pos(outEdges[v], u);
        erase_at(outEdges[v], j);
        if (weighted) erase_at(outEdgeWeights[v], j);
    }
    if (u == v) --selfLoops;
    --m;
    ++ver;
}

edgeweight Graph::weight(node u, node v) const {
    if (u >= z) return 0.0;
    const size_t i = find_pos(outEdges[u], v);
    if (i == outEdges[u].size()) return 0.0;
    return weighted ? outEdgeWeights[u][i] : defaultEdgeWeight;
}

edgeweight Graph::weightedDegree(node u, bool countSelfLoopsTwice) const {
    checked(u);
    if (weighted) {
        edgeweight sum = 0.0;
        const auto &lst = outEdges[u];
        const auto &wl = outEdgeWeights[u];
        for (size_t i = 0; i < lst.size(); ++i)
            sum += (countSelfLoopsTwice && lst[i] == u) ? 2.0 * wl[i] : wl[i];
        return sum;
    }
    count sum = outEdges[u].size();
    if (countSelfLoopsTwice && selfLoops)
        for (node v : outEdges[u]) sum += (v == u);
    return static_cast<edgeweight>(sum);
}

Graph Graph::fromLists(count z_, bool weighted_, bool directed_, const uint8_t *exists_,
                       const uint64_t *outPtr, const int32_t *outIds, const double *outW,
                       const uint64_t *pullPtr, const int32_t *pullIds, const double *pullW,
                       count edges, count loops) {
    Graph G(z_, weighted_, directed_);
    auto fill = [&](std::vector<std::vector<node>> &lists,
                    std::vector<std::vector<edgeweight>> &wl, const uint64_t *ptr,
                    const int32_t *ids, const double *w) {
        for (node u = 0; u < z_; ++u) {
            const uint64_t b = ptr[u], e = ptr[u + 1];
            lists[u].assign(ids + b, ids + e);
            if (weighted_) {
                if (w) wl[u].assign(w + b, w + e);
                else wl[u].assign(e - b, defaultEdgeWeight);
            }
        }
    };
    if (directed_) {
        fill(G.outEdges, G.outEdgeWeights, outPtr, outIds, outW);
        fill(G.inEdges, G.inEdgeWeights, pullPtr, pullIds, pullW);
    } else {
        fill(G.outEdges, G.outEdgeWeights, pullPtr, pullIds, pullW);
    }
    if (exists_)
        for (node u = 0; u < z_; ++u)
            if (!exists_[u]) {
                G.exists[u] = false;
                --G.n;
            }
    G.m = edges;
    G.selfLoops = loops;
    G.ver = 1;
    return G;
}

void Graph::exportLists(bool outgoing, std::vector<uint64_t> &colptr, std::vector<int32_t> &src,
                        std::vector<float> *w32, std::vector<double> *w64) const {
    if (z >= 0x7FFFFFFFull) throw std::runtime_error("graph too large for int32 node ids");
    const bool in = directed && !outgoing;
    const auto &lists = in ? inEdges : outEdges;
    const auto &wlists = in ? inEdgeWeights : outEdgeWeights;
    colptr.assign(z + 1, 0);
    for (node u = 0; u < z; ++u) colptr[u + 1] = colptr[u] + lists[u].size();
    const uint64_t total = colptr[z];
    src.resize(total);
    if (weighted && w32) w32->resize(total);
    if (weighted && w64) w64->resize(total);
#pragma omp parallel for schedule(guided)
    for (int64_t u = 0; u < (int64_t)z; ++u) {
        const uint64_t b = colptr[u];
        const auto &lst = lists[u];
        for (size_t i = 0; i < lst.size(); ++i) src[b + i] = (int32_t)lst[i];
        if (weighted) {
            const auto &wl = wlists[u];
            for (size_t i = 0; i < wl.size(); ++i) {
                if (w32) (*w32)[b + i] = (float)wl[i];
                if (w64) (*w64)[b + i] = wl[i];
            }
        }
    }
}

}  // namespace NetworKit
