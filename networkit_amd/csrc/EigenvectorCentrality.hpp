// C++ drop-ins for NetworKit::EigenvectorCentrality and NetworKit::KatzCentrality on the MI355X
// library (nkpr_power_run).
//
// Interface kept from the reference (kit-parco/networkit v9.0):
//   EigenvectorCentrality  include/networkit/centrality/EigenvectorCentrality.hpp:21-37
//                          (constructor EigenvectorCentrality(G, tol = 1e-8), run())
//   KatzCentrality         include/networkit/centrality/KatzCentrality.hpp:17-55
//                          (constructor KatzCentrality(G, alpha = 0, beta = 0.1, tol = 1e-8),
//                          run(), the public field edgeDirection, enum EdgeDirection)
// run() exports the graph's lists to the device once per graph version (with unit
// out-degrees, so that the device coefficients are the edge weights), then calls
// nkpr_power_run; a non-zero status becomes std::runtime_error.
#pragma once

#include "PageRank.hpp"

namespace NetworKit {

enum EdgeDirection : char { InEdges = 0, OutEdges = 1 };

// The device copy of a graph's in- or out-lists with unit out-degrees, rebuilt when the graph
// or the direction changes.
class DeviceAdjacency {
public:
    ~DeviceAdjacency();
    nkpr_graph *get(const Graph &G, bool outgoing);

private:
    nkpr_graph *dev = nullptr;
    uint64_t version = 0;
    bool out = false;
};

class EigenvectorCentrality final : public Centrality {
    const double tol;

public:
    EigenvectorCentrality(const Graph &G, double tol = 1e-8);
    void run() override;
    const nkpr_stats &lastStats() const { return stats; }

private:
    DeviceAdjacency adj;
    nkpr_stats stats{};
};

class KatzCentrality : public Centrality {
protected:
    const double alpha;  // damping
    const double beta;   // constant centrality amount
    const double tol;    // error tolerance
    static double defaultAlpha(const Graph &G);

public:
    KatzCentrality(const Graph &G, double alpha = 0, double beta = 0.1, double tol = 1e-8);
    void run() override;
    EdgeDirection edgeDirection = EdgeDirection::InEdges;
    const nkpr_stats &lastStats() const { return stats; }

private:
    DeviceAdjacency adj;
    nkpr_stats stats{};
};

}  // namespace NetworKit
