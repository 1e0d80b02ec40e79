#include "EigenvectorCentrality.hpp"

#include <algorithm>
#include <string>
#include <vector>

namespace NetworKit {

DeviceAdjacency::~DeviceAdjacency() {
    if (dev) nkpr_graph_destroy(dev);
}

nkpr_graph *DeviceAdjacency::get(const Graph &G, bool outgoing) {
    const bool lists_out = outgoing && G.isDirected();
    if (dev && version == G.version() && out == lists_out) return dev;
    if (dev) {
        nkpr_graph_destroy(dev);
        dev = nullptr;
    }
    std::vector<uint64_t> colptr;
    std::vector<int32_t> src;
    std::vector<float> w32;
    std::vector<double> w64;
    G.exportLists(lists_out, colptr, src, G.isWeighted() ? &w32 : nullptr,
                  G.isWeighted() ? &w64 : nullptr);
    const count z = G.upperNodeIdBound();
    // unit out-degrees: the device folds w / deg into its coefficients, which are then w
    std::vector<double> ones(z, 1.0);
    std::vector<uint8_t> ex(z, 0);
    G.forNodes([&](node u) { ex[u] = 1; });
    nkpr_csc c{};
    c.z = z;
    c.colptr = colptr.data();
    c.src = src.data();
    c.weight = G.isWeighted() ? w32.data() : nullptr;
    c.weight64 = G.isWeighted() ? w64.data() : nullptr;
    c.out_wdeg = ones.data();
    c.exists = ex.data();
    c.device_ptrs = 0;
    if (nkpr_graph_create(&c, &dev) != NKPR_OK) {
        dev = nullptr;
        throw std::runtime_error(std::string("nkpr_graph_create: ") + nkpr_last_error());
    }
    version = G.version();
    out = lists_out;
    return dev;
}

namespace {
void run_power(nkpr_graph *dev, const nkpr_power_params &p, std::vector<double> &scores,
               count z, nkpr_stats &stats) {
    scores.assign(z, 1.0);
    if (nkpr_power_run(dev, &p, scores.data(), &stats) != NKPR_OK)
        throw std::runtime_error(std::string("nkpr_power_run: ") + nkpr_last_error());
}
}  // namespace

// ---- EigenvectorCentrality (EigenvectorCentrality.cpp:16-66) --------------------------------
EigenvectorCentrality::EigenvectorCentrality(const Graph &G_, double tol_)
    : Centrality(G_, true), tol(tol_) {}

void EigenvectorCentrality::run() {
    nkpr_graph *dev = adj.get(G, false);
    nkpr_power_params p;
    nkpr_power_params_default(&p);
    p.kind = NKPR_EIGENVECTOR;
    p.tol = tol;
    run_power(dev, p, scoreData, G.upperNodeIdBound(), stats);
    hasRun = true;
}

// ---- KatzCentrality (KatzCentrality.cpp:15-76) ----------------------------------------------
double KatzCentrality::defaultAlpha(const Graph &G) {
    // 1 / (1 + GraphTools::maxDegree(G)): the largest unweighted out-degree over all slots
    count maxDeg = 0;
    for (node u = 0; u < G.upperNodeIdBound(); ++u) maxDeg = std::max(maxDeg, G.degreeOut(u));
    return 1. / (1. + static_cast<double>(maxDeg));
}

KatzCentrality::KatzCentrality(const Graph &G_, double alpha_, double beta_, double tol_)
    : Centrality(G_, true), alpha(alpha_ == 0 ? KatzCentrality::defaultAlpha(G_) : alpha_),
      beta(beta_), tol(tol_) {
    if (alpha_ < 0) throw std::runtime_error("Alpha must be non negative.");
}

void KatzCentrality::run() {
    if (edgeDirection != EdgeDirection::InEdges && edgeDirection != EdgeDirection::OutEdges)
        throw std::runtime_error("Unsupported edge direction");
    nkpr_graph *dev = adj.get(G, edgeDirection == EdgeDirection::OutEdges);
    nkpr_power_params p;
    nkpr_power_params_default(&p);
    p.kind = NKPR_KATZ;
    p.alpha = alpha;
    p.beta = beta;
    p.tol = tol;
    run_power(dev, p, scoreData, G.upperNodeIdBound(), stats);
    hasRun = true;
}

}  // namespace NetworKit
