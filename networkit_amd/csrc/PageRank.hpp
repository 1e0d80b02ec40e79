// C++ drop-in for NetworKit::PageRank : Centrality : Algorithm on the MI355X library.
//
// Interface kept from the reference (kit-parco/networkit v9.0):
//   Algorithm   include/networkit/base/Algorithm.hpp:15-57   run / hasFinished / assureFinished
//   Centrality  include/networkit/centrality/Centrality.hpp:20-97, Centrality.cpp:14-75
//   PageRank    include/networkit/centrality/PageRank.hpp:28-75 (ctor defaults, public fields
//               maxIterations and norm, numberOfIterations, maximum)
// Added: the SinkHandling constructor argument (default NoSinkHandling keeps reference
// results), and the fields `precision` and `certifiedStop` that select the mixed-precision
// path of the device library.
//
// run() exports the graph's pull lists to the device once per graph version, then calls
// nkpr_run.  A non-zero status of the C ABI becomes std::runtime_error (or
// Aux::SignalHandling::InterruptException when SIGINT arrived), as the reference's run()
// throws.
#pragma once

#include <atomic>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nkpr.h"
#include "Graph.hpp"

namespace Aux {
namespace SignalHandling {
class InterruptException : public std::exception {
public:
    const char *what() const noexcept override { return "Received a SIGINT"; }
};
}  // namespace SignalHandling

// SIGINT guard in the style of Aux::SignalHandler (networkit/cpp/auxiliary/SignalHandling.cpp):
// the outermost instance installs the handler and restores the previous one on destruction.
class SignalHandler {
public:
    SignalHandler();
    ~SignalHandler();
    void assureRunning();
    bool isRunning();
    static const volatile int32_t *flag();
};
}  // namespace Aux

namespace NetworKit {

class Algorithm {
protected:
    bool hasRun = false;

public:
    Algorithm() = default;
    virtual ~Algorithm() = default;
    virtual void run() = 0;
    bool hasFinished() const noexcept { return hasRun; }
    void assureFinished() const {
        if (!hasRun) throw std::runtime_error("Error, run must be called first");
    }
    virtual std::string toString() const { return "Algorithm"; }
    virtual bool isParallel() const { return true; }
};

class Centrality : public Algorithm {
public:
    Centrality(const Graph &G, bool normalized = false, bool computeEdgeCentrality = false);
    ~Centrality() override = default;
    void run() override = 0;
    virtual const std::vector<double> &scores() const;
    virtual std::vector<double> edgeScores();
    virtual std::vector<std::pair<node, double>> ranking();
    virtual double score(node v);
    virtual double maximum();
    virtual double centralization();

protected:
    const Graph &G;
    std::vector<double> scoreData;
    std::vector<double> edgeScoreData;
    bool normalized;
    bool computeEdgeCentrality;
};

class PageRank final : public Centrality {
public:
    enum Norm { L1Norm, L2Norm };
    enum SinkHandling { NoSinkHandling, DistributeSinks };
    enum Precision { FP64 = NKPR_FP64, Mixed = NKPR_MIXED, FP32 = NKPR_FP32 };

    PageRank(const Graph &G, double damp = 0.85, double tol = 1e-8, bool normalized = false,
             SinkHandling distributeSinks = NoSinkHandling);
    ~PageRank() override;
    PageRank(const PageRank &) = delete;
    PageRank &operator=(const PageRank &) = delete;

    void run() override;
    double maximum() override;
    count numberOfIterations() const {
        assureFinished();
        return iterations;
    }

    count maxIterations = std::numeric_limits<count>::max();
    Norm norm = Norm::L2Norm;

    // device-path knobs (not in the reference)
    Precision precision = Precision::FP64;
    bool certifiedStop = false;  // stop on the certified L1 error bound instead of the step
    bool timeSweeps = false;     // fill lastStats() sweep times with HIP events
    const nkpr_stats &lastStats() const { return stats; }
    const nkpr_graph_info &deviceGraphInfo();

private:
    void ensureDeviceGraph();
    double damp;
    double tol;
    count iterations = 0;
    bool normalizedScores;
    SinkHandling sinkHandling;
    std::atomic<double> max{0.0};
    nkpr_graph *dev = nullptr;
    uint64_t devVersion = 0;
    nkpr_stats stats{};
    nkpr_graph_info info{};
};

}  // namespace NetworKit
