// C++ drop-ins for NetworKit's binary graph reader and writer on the library's nkbg decoder
// (nkpr_nkbg_read / nkpr_nkbg_write).
//
// Interface kept from the reference (kit-parco/networkit v9.0):
//   NetworkitBinaryReader  include/networkit/io/NetworkitBinaryReader.hpp:21-35 (read(path))
//   NetworkitBinaryWriter  include/networkit/io/NetworkitBinaryWriter.hpp:20-45 (constructor
//                          (chunks = 32, weightsType = autoDetect), write(G, path)) and the
//                          enum NetworkitBinaryWeights
// read() builds the Graph from the decoder's flat lists in one pass per list (no per-edge
// insertion); write() exports the Graph's lists in storage order.
#pragma once

#include <cstdint>
#include <string>

#include "../../include/nkpr.h"
#include "Graph.hpp"

namespace NetworKit {

enum class NetworkitBinaryWeights : int {
    none,
    unsignedFormat,
    signedFormat,
    doubleFormat,
    floatFormat,
    autoDetect
};

class NetworkitBinaryReader {
public:
    NetworkitBinaryReader() = default;
    Graph read(const std::string &path);
};

class NetworkitBinaryWriter {
public:
    explicit NetworkitBinaryWriter(uint64_t chunks = 32,
                                   NetworkitBinaryWeights weightsType = NetworkitBinaryWeights::autoDetect)
        : chunks(chunks), weightsType(weightsType) {}
    void write(const Graph &G, const std::string &path);

private:
    uint64_t chunks;
    NetworkitBinaryWeights weightsType;
};

}  // namespace NetworKit
