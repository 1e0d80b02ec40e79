"""Graph readers that produce the node numbering NetworKit's readers produce.

- ``SNAPGraphReader``: the reference's PageRank known answers are stated in the node ids of
  NetworKit's SNAP reader (networkit/cpp/io/SNAPGraphReader.cpp:18-111).
- ``NetworkitBinaryReader`` / ``NetworkitBinaryWriter``: NetworKit's binary format (nkbg002,
  networkit/cpp/io/NetworkitBinaryReader.cpp, NetworkitBinaryWriter.cpp) through the C++
  drop-ins; ``readNetworkitBinaryLists`` decodes a file straight into the flat pull lists the
  device library takes (no Graph in between).
"""
import gzip

import numpy as np


def _open_text(path):
    if str(path).endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path, "r")


def read_snap_edges(path, directed=True, remap_nodes=True):
    """Parse a SNAP edge list the way SNAPGraphReader does and return ``(z, u, v)``.

    - lines starting with '#' and empty lines are skipped (SNAPGraphReader.cpp:79-85);
    - with ``remap_nodes`` ids are renumbered 0, 1, ... in first-seen order, source id of a
      line before its target id (:27-40; both argument orders give the same numbering on the
      reference's wiki-Vote fixture);
    - without remapping, ids are used as given and every id below the largest is a node;
    - an edge already present is dropped (``hasEdge`` check, :43-50).  For an undirected
      graph {u, v} and {v, u} are the same edge.
    """
    rows = []
    with _open_text(path) as f:
        for line in f:
            s = line.strip()
            if not s or s[0] == "#":
                continue
            a, b = s.split()[:2]
            rows.append((int(a), int(b)))
    raw = np.array(rows, dtype=np.int64).reshape(-1, 2)
    if remap_nodes:
        flat = raw.reshape(-1)
        uniq, first = np.unique(flat, return_index=True)
        order = np.argsort(first, kind="stable")
        new_id = np.empty(len(uniq), dtype=np.int64)
        new_id[order] = np.arange(len(uniq), dtype=np.int64)
        mapped = new_id[np.searchsorted(uniq, flat)].reshape(-1, 2)
        z = len(uniq)
    else:
        mapped = raw
        z = int(raw.max()) + 1 if len(raw) else 0
    u, v = mapped[:, 0], mapped[:, 1]
    if directed:
        key = u * max(z, 1) + v
    else:
        key = np.minimum(u, v) * max(z, 1) + np.maximum(u, v)
    _, keep = np.unique(key, return_index=True)
    keep.sort()
    return z, u[keep], v[keep]


class SNAPGraphReader:
    """``SNAPGraphReader(directed).read(path)`` -> Graph (SNAPGraphReader.hpp)."""

    def __init__(self, directed=False, remapNodes=True, nodeCount=0):
        self.directed = directed
        self.remapNodes = remapNodes

    def read(self, path):
        from .graph import Graph
        z, u, v = read_snap_edges(path, self.directed, self.remapNodes)
        G = Graph(z, False, self.directed)
        G.addEdges(u, v)
        return G


class NetworkitBinaryWeights:
    """NetworkitBinaryWriter.hpp: the weight format of a written file."""
    none, unsignedFormat, signedFormat, doubleFormat, floatFormat, autoDetect = range(6)


class NetworkitBinaryReader:
    """``NetworkitBinaryReader().read(path)`` -> Graph (NetworkitBinaryReader.hpp:21-35)."""

    def read(self, path):
        from ._core import _read_nkbg
        return _read_nkbg(str(path))


class NetworkitBinaryWriter:
    """``NetworkitBinaryWriter(chunks=32, weightsType=autoDetect).write(G, path)``
    (NetworkitBinaryWriter.hpp:20-45)."""

    def __init__(self, chunks=32, weightsType=NetworkitBinaryWeights.autoDetect):
        self.chunks = chunks
        self.weightsType = weightsType

    def write(self, G, path):
        from ._core import _write_nkbg
        _write_nkbg(G, str(path), self.chunks, self.weightsType)


def readNetworkitBinaryLists(path, with_out_lists=False):
    """The pull lists of an nkbg002 file as numpy arrays (colptr, src, weight, exists, ...),
    ready for ``capi.DeviceGraph`` (see ``capi.read_nkbg``)."""
    from .capi import read_nkbg
    return read_nkbg(path, with_out_lists)
