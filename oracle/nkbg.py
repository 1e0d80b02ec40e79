"""NetworKit binary graph format (nkbg002) -- TEST INFRASTRUCTURE ONLY.

A plain-Python restatement of the format as NetworKit writes and reads it, used by the tests
to write input files and to check the library's reader:

  varint / zigzag coding      include/networkit/io/NetworkitBinaryGraph.hpp:48-141
  writer                      networkit/cpp/io/NetworkitBinaryWriter.cpp:19-350
  reader                      networkit/cpp/io/NetworkitBinaryReader.cpp:19-237
  layout notes                networkit/cpp/io/NetworkitBinaryGraph.md

Layout (all integers little-endian):
  header       magic "nkbg002\\0", checksum, features (bit 0 directed, bits 1-3 weight format:
               0 none, 1 unsigned varint, 2 signed (zigzag) varint, 3 double, 4 float), nodes,
               chunks, and the offsets of the base data, adjacency lists, transposed lists,
               weights and transposed weights (10 x 8 bytes)
  base data    one flag byte per node (bit 0 set: the node exists), then the first node of
               chunks 1 .. chunks-1 (uint64 each)
  adjacency    per-chunk offsets of chunks 1 .. chunks-1 (uint64), the number of list entries
               (uint64), then per node: varint count, varint neighbours
  transposed   the same for the transposed lists
  weights      per-chunk offsets, then the weights of the adjacency entries in order
  t. weights   per-chunk offsets, then the weights of the transposed entries
An undirected graph stores at node u the neighbours v <= u in the adjacency lists and
v >= u in the transposed lists (a self-loop in both); the reader rebuilds u's adjacency as the
first list followed by the second without its self-loops.  A directed graph stores the
out-lists and the in-lists.

Parity: the varint coding is pinned by the worked examples of NetworkitBinaryGraph.hpp:55-72
and by the round trip over every bit width (IOGTest.cpp:987-1043); zigzag by its round trip
and bound (IOGTest.cpp:1045-1064); the writer and reader by hand-built byte strings of small
graphs and by the reference's round-trip checks (IOGTest.cpp:822-985) on its own input files.
"""
import struct

import numpy as np

MAGIC = b"nkbg002\0"
NONE, VARINT, SIGNED_VARINT, DOUBLE, FLOAT = 0, 1, 2, 3, 4


def varint_encode(value):
    """NetworkitBinaryGraph.hpp:75-105: the lowest set bit of the first byte gives the number
    of data bytes that follow (0 to 7); a first byte of 0 means 8 data bytes.  The first byte's
    remaining bits hold the lowest bits of the value, the data bytes the rest, little-endian."""
    value = int(value)
    if value == 0:
        return bytes([1])
    if value >= 1 << 56:
        return bytes([0]) + value.to_bytes(8, "little")
    nbytes = (value.bit_length() - 1) // 7
    head = ((1 << nbytes) | (value << (nbytes + 1))) & 0xFF
    rest = value >> (7 - nbytes)
    return bytes([head]) + rest.to_bytes(nbytes, "little") if nbytes else bytes([head])


def varint_decode(buf, pos):
    """NetworkitBinaryGraph.hpp:108-126.  Returns (value, bytes consumed)."""
    head = buf[pos]
    if head == 0:
        return int.from_bytes(bytes(buf[pos + 1:pos + 9]), "little"), 9
    n = (head & -head).bit_length() - 1
    value = head >> (n + 1)
    bits = 7 - n
    for i in range(n):
        value |= buf[pos + 1 + i] << bits
        bits += 8
    return value, n + 1


def zigzag_encode(value):
    """NetworkitBinaryGraph.hpp:132-134 (64-bit)."""
    value = int(value)
    return ((value << 1) ^ (-1 if value < 0 else 0)) & ((1 << 64) - 1)


def zigzag_decode(value):
    """NetworkitBinaryGraph.hpp:137-139."""
    value = int(value)
    return (value >> 1) ^ -(value & 1)


class ListGraph:
    """NetworKit's adjacency as plain lists: out[u] = [(v, w), ...] in storage order, and for a
    directed graph inn[u] likewise (an undirected edge appears at both endpoints, a self-loop
    once, Graph.cpp:549-597).  exists[u] False for deleted slots."""

    def __init__(self, n, weighted=False, directed=False):
        self.n = n
        self.weighted = weighted
        self.directed = directed
        self.exists = [True] * n
        self.out = [[] for _ in range(n)]
        self.inn = [[] for _ in range(n)] if directed else None

    @classmethod
    def from_edges(cls, n, u, v, w=None, directed=False):
        g = cls(n, w is not None, directed)
        for i in range(len(u)):
            a, b = int(u[i]), int(v[i])
            wt = 1.0 if w is None else float(w[i])
            g.out[a].append((b, wt))
            if directed:
                g.inn[b].append((a, wt))
            elif a != b:
                g.out[b].append((a, wt))
        return g

    def number_of_edges(self):
        if self.directed:
            return sum(len(x) for x in self.out)
        loops = sum(1 for u in range(self.n) for v, _ in self.out[u] if v == u)
        return (sum(len(x) for x in self.out) - loops) // 2 + loops

    def self_loops(self):
        return sum(1 for u in range(self.n) for v, _ in self.out[u] if v == u)

    def pull_lists(self):
        """Rows PageRank pulls over: in-lists (directed) or the adjacency (undirected)."""
        return self.inn if self.directed else self.out


def _weight_format(g, kind):
    if kind != "auto":
        return kind
    if not g.weighted:
        return NONE
    ws = [w for lst in g.out for _, w in lst]
    unsigned = all(w >= 0 for w in ws)
    ints = all(w == int(w) for w in ws if abs(w) < 2 ** 63)
    floats = all(float(np.float32(w)) == w for w in ws)
    if ints:
        return VARINT if unsigned else SIGNED_VARINT
    return FLOAT if floats else DOUBLE


def _weight_bytes(w, fmt):
    if fmt == VARINT:
        return varint_encode(int(w))
    if fmt == SIGNED_VARINT:
        return varint_encode(zigzag_encode(int(w)))
    if fmt == DOUBLE:
        return struct.pack("<d", w)
    if fmt == FLOAT:
        return struct.pack("<f", w)
    return b""


def write(g, chunks=32, weights="auto"):
    """NetworkitBinaryWriter::write (NetworkitBinaryWriter.cpp:19-350) of a graph without
    deleted nodes; returns the file's bytes."""
    assert all(g.exists), "the reference writer numbers the nodes 0 .. n-1"
    fmt = _weight_format(g, weights)
    nodes = g.n
    chunks = min(chunks, nodes)
    first = [0] + [c * (nodes // chunks) for c in range(1, chunks)] + [nodes]

    def lists(u):
        if g.directed:
            return g.out[u], g.inn[u]
        return ([(v, w) for v, w in g.out[u] if v <= u], [(v, w) for v, w in g.out[u] if v >= u])

    adj, tr, adjw, trw = [], [], [], []
    adj_off, tr_off, adjw_off, trw_off = [], [], [], []
    a_size = t_size = 0
    for c in range(chunks):
        for u in range(first[c], first[c + 1]):
            o, i = lists(u)
            adj.append(varint_encode(len(o)) + b"".join(varint_encode(v) for v, _ in o))
            tr.append(varint_encode(len(i)) + b"".join(varint_encode(v) for v, _ in i))
            adjw.append(b"".join(_weight_bytes(w, fmt) for _, w in o))
            trw.append(b"".join(_weight_bytes(w, fmt) for _, w in i))
            a_size += len(o)
            t_size += len(i)
        adj_off.append(sum(len(x) for x in adj))
        tr_off.append(sum(len(x) for x in tr))
        adjw_off.append(sum(len(x) for x in adjw))
        trw_off.append(sum(len(x) for x in trw))
    u64 = lambda x: struct.pack("<Q", x)  # noqa: E731
    base = 80
    o_adj = base + nodes + (chunks - 1) * 8
    o_tr = o_adj + (chunks - 1) * 8 + 8 + adj_off[-1]
    if fmt != NONE:
        o_w = o_tr + (chunks - 1) * 8 + 8 + tr_off[-1]
        o_tw = o_w + (chunks - 1) * 8 + adjw_off[-1]
    else:
        o_w = o_tw = 0
    features = (1 if g.directed else 0) | ((fmt << 1) & 0xE)
    out = [MAGIC, u64(0), u64(features), u64(nodes), u64(chunks), u64(base), u64(o_adj), u64(o_tr),
           u64(o_w), u64(o_tw)]
    out.append(bytes([1] * nodes))
    out += [u64(first[c]) for c in range(1, chunks)]
    out += [u64(adj_off[c - 1]) for c in range(1, chunks)] + [u64(a_size)] + adj
    out += [u64(tr_off[c - 1]) for c in range(1, chunks)] + [u64(t_size)] + tr
    out += [u64(adjw_off[c - 1]) for c in range(1, chunks)] + adjw
    out += [u64(trw_off[c - 1]) for c in range(1, chunks)] + trw
    return b"".join(out)


def read(data):
    """NetworkitBinaryReader::read (NetworkitBinaryReader.cpp:19-237) of a file's bytes.
    Returns (ListGraph, number of edges, number of self-loops)."""
    buf = memoryview(data)
    if bytes(buf[:8]) != MAGIC:
        raise ValueError("Reader expected another magic value")
    q = struct.unpack_from("<9Q", buf, 8)
    _, features, nodes, chunks, o_base, o_adj, o_tr, o_w, o_tw = q
    directed = bool(features & 1)
    fmt = (features & 0xE) >> 1
    g = ListGraph(nodes, fmt != NONE, directed)
    for i in range(nodes):
        if not buf[o_base + i] & 1:
            g.exists[i] = False
    first = [0] + [struct.unpack_from("<Q", buf, o_base + nodes + 8 * k)[0]
                   for k in range(chunks - 1)] + [nodes]
    adj_size = struct.unpack_from("<Q", buf, o_adj + (chunks - 1) * 8)[0]
    loops = 0

    def weight(pos):
        if fmt == VARINT:
            v, k = varint_decode(buf, pos)
            return float(v), k
        if fmt == SIGNED_VARINT:
            v, k = varint_decode(buf, pos)
            return float(zigzag_decode(v)), k
        if fmt == DOUBLE:
            return struct.unpack_from("<d", buf, pos)[0], 8
        if fmt == FLOAT:
            return float(struct.unpack_from("<f", buf, pos)[0]), 4
        return 1.0, 0

    for c in range(chunks):
        def start(sec, extra):
            off = struct.unpack_from("<Q", buf, sec + (c - 1) * 8)[0] if first[c] else 0
            return sec + off + (chunks - 1) * 8 + extra
        pa, pt = start(o_adj, 8), start(o_tr, 8)
        pw, ptw = start(o_w, 0), start(o_tw, 0)
        for u in range(first[c], first[c + 1]):
            nout, k = varint_decode(buf, pa)
            pa += k
            nin, k = varint_decode(buf, pt)
            pt += k
            for _ in range(nout):
                v, k = varint_decode(buf, pa)
                pa += k
                w, k = weight(pw)
                pw += k
                g.out[u].append((v, w))
                loops += v == u
            for _ in range(nin):
                v, k = varint_decode(buf, pt)
                pt += k
                w, k = weight(ptw)
                ptw += k
                if directed:
                    g.inn[u].append((v, w))
                elif v != u:
                    g.out[u].append((v, w))
    return g, adj_size, loops
