#include "NetworkitBinaryIO.hpp"

#include <stdexcept>
#include <vector>

namespace NetworKit {

namespace {
struct Owned {
    nkpr_nkbg g{};
    ~Owned() { nkpr_nkbg_free(&g); }
};
}  // namespace

Graph NetworkitBinaryReader::read(const std::string &path) {
    Owned o;
    if (nkpr_nkbg_read(path.c_str(), 1, &o.g) != NKPR_OK) throw std::runtime_error(nkpr_last_error());
    const nkpr_nkbg &g = o.g;
    return Graph::fromLists(g.nodes, g.weighted != 0, g.directed != 0, g.exists, g.out_ptr,
                            g.out_dst, g.out_weight, g.colptr, g.src, g.weight, g.edges,
                            g.self_loops);
}

void NetworkitBinaryWriter::write(const Graph &G, const std::string &path) {
    std::vector<uint64_t> optr, iptr;
    std::vector<int32_t> oids, iids;
    std::vector<double> ow, iw;
    G.exportOutLists(optr, oids, G.isWeighted() ? &ow : nullptr);
    if (G.isDirected()) G.exportInLists(iptr, iids, nullptr, G.isWeighted() ? &iw : nullptr);
    std::vector<uint8_t> ex(G.upperNodeIdBound());
    for (node u = 0; u < G.upperNodeIdBound(); ++u) ex[u] = G.hasNode(u) ? 1 : 0;
    nkpr_nkbg g{};
    g.nodes = G.upperNodeIdBound();
    g.edges = G.numberOfEdges();
    g.self_loops = G.numberOfSelfLoops();
    g.directed = G.isDirected();
    g.weighted = G.isWeighted();
    g.out_ptr = optr.data();
    g.out_dst = oids.data();
    g.out_weight = G.isWeighted() ? ow.data() : nullptr;
    if (G.isDirected()) {
        g.colptr = iptr.data();
        g.src = iids.data();
        g.weight = G.isWeighted() ? iw.data() : nullptr;
    }
    g.exists = ex.data();
    int fmt;
    switch (weightsType) {
    case NetworkitBinaryWeights::none: fmt = 0; break;
    case NetworkitBinaryWeights::unsignedFormat: fmt = 1; break;
    case NetworkitBinaryWeights::signedFormat: fmt = 2; break;
    case NetworkitBinaryWeights::doubleFormat: fmt = 3; break;
    case NetworkitBinaryWeights::floatFormat: fmt = 4; break;
    default: fmt = -1; break;
    }
    if (!G.isWeighted()) fmt = 0;
    if (nkpr_nkbg_write(path.c_str(), &g, chunks, fmt) != NKPR_OK)
        throw std::runtime_error(nkpr_last_error());
}

}  // namespace NetworKit
