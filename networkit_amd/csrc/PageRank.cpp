#include "PageRank.hpp"

#include <algorithm>
#include <csignal>
#include <mutex>

namespace Aux {
namespace {
volatile int32_t g_sigint = 0;
std::mutex g_sig_mutex;
SignalHandler *g_root = nullptr;
void (*g_prev)(int) = SIG_DFL;
void on_sigint(int sig) {
    if (sig == SIGINT) g_sigint = 1;
}
}  // namespace

SignalHandler::SignalHandler() {
    std::lock_guard<std::mutex> lk(g_sig_mutex);
    if (!g_root) {
        g_root = this;
        g_prev = std::signal(SIGINT, on_sigint);
    }
}

SignalHandler::~SignalHandler() {
    std::lock_guard<std::mutex> lk(g_sig_mutex);
    if (g_root == this) {
        g_root = nullptr;
        g_sigint = 0;
        std::signal(SIGINT, g_prev == SIG_ERR ? SIG_DFL : g_prev);
    }
}

void SignalHandler::assureRunning() {
    if (g_sigint) throw SignalHandling::InterruptException();
}

bool SignalHandler::isRunning() { return !g_sigint; }

const volatile int32_t *SignalHandler::flag() { return &g_sigint; }
}  // namespace Aux

namespace NetworKit {

// ---- Centrality (Centrality.cpp:14-75) --------------------------------------------------------
Centrality::Centrality(const Graph &G_, bool normalized_, bool computeEdgeCentrality_)
    : G(G_), normalized(normalized_), computeEdgeCentrality(computeEdgeCentrality_) {
    if (computeEdgeCentrality_)
        throw std::runtime_error("edge centralities are not provided by this library");
}

const std::vector<double> &Centrality::scores() const {
    assureFinished();
    return scoreData;
}

std::vector<double> Centrality::edgeScores() {
    assureFinished();
    return edgeScoreData;
}

double Centrality::score(node v) {
    assureFinished();
    return scoreData.at(v);
}

std::vector<std::pair<node, double>> Centrality::ranking() {
    assureFinished();
    const count z = scoreData.size();
    std::vector<uint8_t> ex(z);
    const auto &flags = G.existsFlags();
    for (count u = 0; u < z; ++u) ex[u] = (u < flags.size() && flags[u]) ? 1 : 0;
    std::vector<int64_t> ids(z);
    std::vector<double> vals(z);
    uint64_t k = 0;
    if (nkpr_ranking(z, scoreData.data(), ex.data(), ids.data(), vals.data(), &k) != NKPR_OK)
        throw std::runtime_error(std::string("nkpr_ranking: ") + nkpr_last_error());
    std::vector<std::pair<node, double>> out;
    out.reserve(k);
    for (uint64_t i = 0; i < k; ++i) out.emplace_back((node)ids[i], vals[i]);
    return out;
}

double Centrality::maximum() {
    throw std::runtime_error("Not implemented: Compute the maximum centrality score in the "
                             "respective centrality subclass.");
}

double Centrality::centralization() {
    assureFinished();
    double centerScore = 0.0;
    G.forNodes([&](node v) {
        if (scoreData[v] > centerScore) centerScore = scoreData[v];
    });
    const double maxScore = maximum();
    double diff1 = 0.0, diff2 = 0.0;
    G.forNodes([&](node v) {
        diff1 += (centerScore - scoreData[v]);
        diff2 += (maxScore - scoreData[v]);
    });
    return diff1 / diff2;
}

// ---- PageRank ---------------------------------------------------------------------------------
PageRank::PageRank(const Graph &G_, double damp_, double tol_, bool normalized_,
                   SinkHandling distributeSinks)
    : Centrality(G_, true), damp(damp_), tol(tol_), normalizedScores(normalized_),
      sinkHandling(distributeSinks) {}

PageRank::~PageRank() {
    if (dev) nkpr_graph_destroy(dev);
}

void PageRank::ensureDeviceGraph() {
    if (dev && devVersion == G.version()) return;
    if (dev) {
        nkpr_graph_destroy(dev);
        dev = nullptr;
    }
    std::vector<uint64_t> colptr;
    std::vector<int32_t> src;
    std::vector<float> w32;
    std::vector<double> w64;
    G.exportInLists(colptr, src, G.isWeighted() ? &w32 : nullptr,
                    G.isWeighted() ? &w64 : nullptr);
    const count z = G.upperNodeIdBound();
    // out-degrees exactly as the reference computes them (PageRank.cpp:29-30)
    std::vector<double> wdeg(z, 0.0);
    std::vector<uint8_t> ex(z, 0);
    G.forNodes([&](node u) {
        wdeg[u] = G.weightedDegree(u);
        ex[u] = 1;
    });
    nkpr_csc c{};
    c.z = z;
    c.colptr = colptr.data();
    c.src = src.data();
    c.weight = G.isWeighted() ? w32.data() : nullptr;
    c.weight64 = G.isWeighted() ? w64.data() : nullptr;
    c.out_wdeg = wdeg.data();
    c.exists = ex.data();
    c.device_ptrs = 0;
    if (nkpr_graph_create(&c, &dev) != NKPR_OK) {
        dev = nullptr;
        throw std::runtime_error(std::string("nkpr_graph_create: ") + nkpr_last_error());
    }
    devVersion = G.version();
    nkpr_graph_get_info(dev, &info);
}

const nkpr_graph_info &PageRank::deviceGraphInfo() {
    ensureDeviceGraph();
    return info;
}

void PageRank::run() {
    Aux::SignalHandler handler;
    const count n = G.numberOfNodes();
    const count z = G.upperNodeIdBound();
    // PageRank.cpp:26: the previous scores stay as the start vector of a re-run
    scoreData.resize(z, 1.0 / static_cast<double>(n));
    handler.assureRunning();
    ensureDeviceGraph();
    nkpr_params p;
    nkpr_params_default(&p);
    p.damp = damp;
    p.tol = tol;
    p.max_iterations = maxIterations;
    p.norm = (norm == L1Norm) ? NKPR_L1 : NKPR_L2;
    p.sink_handling = (sinkHandling == DistributeSinks) ? NKPR_SINK_DISTRIBUTE : NKPR_SINK_NONE;
    p.normalized = normalizedScores ? 1 : 0;
    p.precision = static_cast<int32_t>(precision);
    p.stop = (certifiedStop || precision == Mixed) ? NKPR_STOP_L1_ERROR : NKPR_STOP_STEP;
    p.timing = timeSweeps ? 1 : 0;
    p.interrupt = Aux::SignalHandler::flag();
    const int rc = nkpr_run(dev, &p, scoreData.data(), &stats);
    if (rc == NKPR_ERR_INTERRUPTED) throw Aux::SignalHandling::InterruptException();
    if (rc != NKPR_OK) throw std::runtime_error(std::string("nkpr_run: ") + nkpr_last_error());
    handler.assureRunning();
    iterations = stats.iterations;
    max.store(stats.maximum, std::memory_order_relaxed);
    hasRun = true;
}

double PageRank::maximum() { return max.load(std::memory_order_relaxed); }

}  // namespace NetworKit
