// C++ drop-ins for NetworKit's ApproximatePageRank and PageRankNibble on the library's
// frontier push (nkpr_apr_run).
//
// Interface kept from the reference (kit-parco/networkit v9.0):
//   ApproximatePageRank         include/networkit/scd/ApproximatePageRank.hpp:17-55
//                               (constructor (G, alpha, epsilon = 1e-12), run(seeds), run(seed))
//   SelectiveCommunityDetector  include/networkit/scd/SelectiveCommunityDetector.hpp
//                               (run(seeds) -> seed -> community, expandOneCommunity)
//   PageRankNibble              include/networkit/scd/PageRankNibble.hpp (constructor
//                               (G, alpha, epsilon)), expandOneCommunity(seeds)
// The pushes run on the device; PageRankNibble's sweep over the result runs on the host in
// the reference's order of operations (PageRankNibble.cpp:28-89), with equal scores ordered
// by ascending id (the reference's sort is not stable).
#pragma once

#include <map>
#include <set>
#include <utility>
#include <vector>

#include "../../include/nkpr.h"
#include "Graph.hpp"

namespace NetworKit {

class ApproximatePageRank final {
public:
    ApproximatePageRank(const Graph &g, double alpha, double epsilon = 1e-12);
    ~ApproximatePageRank();
    ApproximatePageRank(const ApproximatePageRank &) = delete;
    ApproximatePageRank &operator=(const ApproximatePageRank &) = delete;
    std::vector<std::pair<node, double>> run(const std::set<node> &seeds);
    std::vector<std::pair<node, double>> run(node seed);
    const nkpr_apr_stats &lastStats() const { return stats; }

private:
    const Graph *g;
    double alpha;
    double eps;
    nkpr_apr *dev = nullptr;
    uint64_t devVersion = 0;
    nkpr_apr_stats stats{};
};

class SelectiveCommunityDetector {
public:
    explicit SelectiveCommunityDetector(const Graph &g) : g(&g) {}
    virtual ~SelectiveCommunityDetector() = default;
    virtual std::map<node, std::set<node>> run(const std::set<node> &seeds);
    virtual std::set<node> expandOneCommunity(node seed);
    virtual std::set<node> expandOneCommunity(const std::set<node> &seeds) = 0;

protected:
    const Graph *g;
};

class PageRankNibble : public SelectiveCommunityDetector {
public:
    PageRankNibble(const Graph &g, double alpha, double epsilon);
    std::set<node> expandOneCommunity(const std::set<node> &seeds) override;
    using SelectiveCommunityDetector::expandOneCommunity;

private:
    std::set<node> bestSweepSet(std::vector<std::pair<node, double>> &pr);
    double alpha;
    double epsilon;
};

}  // namespace NetworKit
