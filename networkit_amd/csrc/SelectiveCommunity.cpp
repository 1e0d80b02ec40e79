#include "SelectiveCommunity.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <unordered_set>

namespace NetworKit {

namespace {
void check(int rc) {
    if (rc != NKPR_OK) throw std::runtime_error(nkpr_last_error());
}
}  // namespace

ApproximatePageRank::ApproximatePageRank(const Graph &g_, double alpha_, double epsilon)
    : g(&g_), alpha(alpha_), eps(epsilon) {}

ApproximatePageRank::~ApproximatePageRank() {
    if (dev) nkpr_apr_destroy(dev);
}

std::vector<std::pair<node, double>> ApproximatePageRank::run(const std::set<node> &seeds) {
    if (!dev || devVersion != g->version()) {
        if (dev) check(nkpr_apr_destroy(dev));
        dev = nullptr;
        std::vector<uint64_t> ptr;
        std::vector<int32_t> ids;
        std::vector<double> w;
        g->exportOutLists(ptr, ids, g->isWeighted() ? &w : nullptr);
        check(nkpr_apr_create(g->upperNodeIdBound(), ptr.data(), ids.data(),
                              g->isWeighted() ? w.data() : nullptr, 0, &dev));
        devVersion = g->version();
    }
    std::vector<uint64_t> sd(seeds.begin(), seeds.end());
    const uint64_t z = g->upperNodeIdBound();
    std::vector<int64_t> idsOut(z);
    std::vector<double> p(z);
    uint64_t k = 0;
    check(nkpr_apr_run(dev, alpha, eps, sd.data(), sd.size(), &k, idsOut.data(), p.data(),
                       nullptr, &stats));
    std::vector<std::pair<node, double>> pr;
    pr.reserve(k);
    for (uint64_t i = 0; i < k; ++i) pr.emplace_back((node)idsOut[i], p[i]);
    return pr;
}

std::vector<std::pair<node, double>> ApproximatePageRank::run(node seed) {
    return run(std::set<node>{seed});
}

std::map<node, std::set<node>> SelectiveCommunityDetector::run(const std::set<node> &seeds) {
    // SelectiveCommunityDetector.cpp:15-23
    std::map<node, std::set<node>> result;
    for (node s : seeds) result[s] = expandOneCommunity(s);
    return result;
}

std::set<node> SelectiveCommunityDetector::expandOneCommunity(node seed) {
    return expandOneCommunity(std::set<node>{seed});
}

PageRankNibble::PageRankNibble(const Graph &g_, double alpha_, double epsilon_)
    : SelectiveCommunityDetector(g_), alpha(alpha_), epsilon(epsilon_) {}

std::set<node> PageRankNibble::bestSweepSet(std::vector<std::pair<node, double>> &pr) {
    // PageRankNibble.cpp:28-89: order by p / weightedDegree(u, true), descending (ties by id)
    for (auto &e : pr) e.second = e.second / g->weightedDegree(e.first, true);
    std::stable_sort(pr.begin(), pr.end(),
                     [](const std::pair<node, double> &a, const std::pair<node, double> &b) {
                         const bool na = std::isnan(a.second), nb = std::isnan(b.second);
                         if (na || nb) return !na && nb;
                         return a.second > b.second;
                     });
    double bestCond = std::numeric_limits<double>::max();
    double cut = 0.0, volume = 0.0;
    size_t bestIndex = 0;
    std::unordered_set<node> within;
    std::vector<node> current;
    const double totalVolume = g->totalEdgeWeight() * 2;
    for (const auto &e : pr) {
        const node v = e.first;
        double wDegree = 0.0;
        g->forNeighborsOf(v, [&](node neigh, edgeweight w) {
            wDegree += w;
            if (within.find(neigh) == within.end()) cut += w;
            else cut -= w;
        });
        volume += wDegree;
        current.push_back(v);
        within.insert(v);
        const double cond = cut / std::min(volume, totalVolume - volume);
        if (cond < bestCond && current.size() < g->numberOfNodes()) {
            bestCond = cond;
            bestIndex = current.size();
        }
    }
    return std::set<node>(current.begin(), current.begin() + bestIndex);
}

std::set<node> PageRankNibble::expandOneCommunity(const std::set<node> &seeds) {
    ApproximatePageRank apr(*g, alpha, epsilon);
    auto pr = apr.run(seeds);
    return bestSweepSet(pr);
}

}  // namespace NetworKit
