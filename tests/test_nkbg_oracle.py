"""The nkbg oracle (oracle/nkbg.py) against what fixes the format: the worked examples and
round-trip checks of the varint and zigzag codes (NetworkitBinaryGraph.hpp:55-72,
IOGTest.cpp:987-1064), hand-built byte strings of small graphs written from the writer's and
the reader's statements (NetworkitBinaryWriter.cpp, NetworkitBinaryReader.cpp), and the
reference's own round-trip checks (IOGTest.cpp:822-985) on its input files."""
import struct

import numpy as np
import pytest

from nkbg_inputs import konect, metis, wiki_vote_edges
from oracle import nkbg


def test_varint_worked_examples():
    # x = 0b0GFE'DCBA with G = 1 (7 bits): one byte GFEDCBA|1
    assert nkbg.varint_encode(0b1000101) == bytes([0b10001011])
    # x = 0bHGFE'DCBA with H = 1 (8 bits): FEDCBA|10, then 000000|HG
    assert nkbg.varint_encode(0b11000101) == bytes([0b00010110, 0b11])
    assert nkbg.varint_encode(0) == bytes([1])
    assert nkbg.varint_encode(1 << 56) == bytes([0]) + (1 << 56).to_bytes(8, "little")
    assert nkbg.varint_decode(bytes([0b00010110, 0b11]), 0) == (0b11000101, 2)


def test_varint_round_trip_every_width():
    rng = np.random.default_rng(1)
    seen = 0
    for bits in range(64):
        lo = 0 if bits == 0 else 1 << bits
        hi = 0 if bits == 0 else 2 * lo - 1
        for i in range(200):
            x = lo if i == 0 else hi if i == 1 else int(rng.integers(lo, hi, endpoint=True,
                                                                    dtype=np.uint64))
            b = nkbg.varint_encode(x)
            assert b[-1] > 0 or len(b) == 1
            assert nkbg.varint_decode(b + bytes(range(10)), 0) == (x, len(b))
            assert len(b) == (9 if x >= 1 << 56 else 1 + (max(x.bit_length(), 1) - 1) // 7)
            seen |= x
    assert seen == (1 << 64) - 1


def test_zigzag():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x = int(rng.integers(0, (1 << 63) - 1, dtype=np.int64))
        for v in (x, -x):
            e = nkbg.zigzag_encode(v)
            assert nkbg.zigzag_decode(e) == v
            assert e <= 2 * abs(v)
    assert [nkbg.zigzag_encode(v) for v in (0, -1, 1, -2, 2)] == [0, 1, 2, 3, 4]


def _header(features, nodes, chunks, o_adj, o_tr, o_w=0, o_tw=0):
    return b"nkbg002\0" + struct.pack("<9Q", 0, features, nodes, chunks, 80, o_adj, o_tr, o_w,
                                      o_tw)


def test_hand_built_directed_unweighted():
    """3 nodes, edges 0->1, 0->2, 2->0, one chunk: the bytes the writer produces."""
    g = nkbg.ListGraph.from_edges(3, [0, 0, 2], [1, 2, 0], directed=True)
    adj = bytes([0b101, 0b11, 0b101, 0b1, 0b11, 0b1])          # 2:[1,2]  0:[]  1:[0]
    tr = bytes([0b11, 0b101, 0b11, 0b1, 0b11, 0b1])            # 1:[2]  1:[0]  1:[0]
    o_adj = 80 + 3
    o_tr = o_adj + 8 + len(adj)
    want = (_header(1, 3, 1, o_adj, o_tr) + bytes([1, 1, 1]) + struct.pack("<Q", 3) + adj +
            struct.pack("<Q", 3) + tr)
    assert nkbg.write(g, chunks=1) == want
    back, edges, loops = nkbg.read(want)
    assert edges == 3 and loops == 0
    assert back.out == [[(1, 1.0), (2, 1.0)], [], [(0, 1.0)]]
    assert back.inn == [[(2, 1.0)], [(0, 1.0)], [(0, 1.0)]]


def test_hand_built_undirected_float_weights_two_chunks():
    """4 nodes, edges {0,1} w 0.5, {1,1} w 2 (self-loop), {3,1} w 0.25; chunks {0,1}, {2,3}.
    Node u keeps v <= u in the adjacency lists and v >= u in the transposed lists."""
    g = nkbg.ListGraph.from_edges(4, [0, 1, 3], [1, 1, 1], [0.5, 2.0, 0.25])
    # adjacency: 0:[0]? no (v<=0: none) -> count 0; 1:[0, 1]; 2:[]; 3:[1]
    adj_c0 = bytes([0b1]) + bytes([0b101, 0b1, 0b11])
    adj_c1 = bytes([0b1]) + bytes([0b11, 0b11])
    # transposed: 0:[1]; 1:[1, 3]; 2:[]; 3:[]
    tr_c0 = bytes([0b11, 0b11]) + bytes([0b101, 0b11, 0b111])
    tr_c1 = bytes([0b1, 0b1])
    f = lambda *w: b"".join(struct.pack("<f", x) for x in w)  # noqa: E731
    aw_c0, aw_c1 = f(0.5, 2.0), f(0.25)
    tw_c0, tw_c1 = f(0.5, 2.0, 0.25), b""
    o_adj = 80 + 4 + 8
    o_tr = o_adj + 8 + 8 + len(adj_c0 + adj_c1)
    o_w = o_tr + 8 + 8 + len(tr_c0 + tr_c1)
    o_tw = o_w + 8 + len(aw_c0 + aw_c1)
    feat = 0 | (nkbg.FLOAT << 1)
    want = (_header(feat, 4, 2, o_adj, o_tr, o_w, o_tw) + bytes([1] * 4) + struct.pack("<Q", 2) +
            struct.pack("<QQ", len(adj_c0), 3) + adj_c0 + adj_c1 +
            struct.pack("<QQ", len(tr_c0), 3) + tr_c0 + tr_c1 +
            struct.pack("<Q", len(aw_c0)) + aw_c0 + aw_c1 +
            struct.pack("<Q", len(tw_c0)) + tw_c0 + tw_c1)
    assert nkbg.write(g, chunks=2) == want
    back, edges, loops = nkbg.read(want)
    assert edges == 3 and loops == 1
    # u's adjacency: adjacency part, then the transposed part without self-loops
    assert back.out == [[(1, 0.5)], [(0, 0.5), (1, 2.0), (3, 0.25)], [], [(1, 0.25)]]


def _round_trip(g, **kw):
    back, edges, loops = nkbg.read(nkbg.write(g, **kw))
    assert back.directed == g.directed and back.weighted == g.weighted
    assert back.n == g.n and edges == g.number_of_edges() and loops == g.self_loops()
    for u in range(g.n):
        assert sorted(back.out[u]) == sorted(g.out[u])
        if g.directed:
            assert sorted(back.inn[u]) == sorted(g.inn[u])
    return back


@pytest.mark.parametrize("name", ["tiny_01.graph", "jazz.graph"])
def test_round_trip_metis_inputs(name):
    n, u, v = metis(name)
    _round_trip(nkbg.ListGraph.from_edges(n, u, v))


def test_round_trip_konect_and_wiki():
    n, u, v, w = konect("foodweb-baydry.konect")
    g = _round_trip(nkbg.ListGraph.from_edges(n, u, v, w, directed=True))
    assert g.weighted
    z, u, v = wiki_vote_edges()
    _round_trip(nkbg.ListGraph.from_edges(z, u, v, directed=True))


@pytest.mark.parametrize("ws,fmt", [([-1, 0, 1, 2, 3, 4, 5, 6, 7], nkbg.SIGNED_VARINT),
                                    ([987.654 + i for i in range(9)], nkbg.FLOAT),
                                    ([0.1 * i for i in range(9)], nkbg.DOUBLE),
                                    ([3, 0, 7, 1 << 40, 5, 6, 7, 8, 9], nkbg.VARINT)])
def test_weight_formats(ws, fmt):
    # IOGTest.cpp:907-951: a path of 10 nodes with signed / float weights
    ws = [float(np.float32(x)) if fmt == nkbg.FLOAT else float(x) for x in ws]
    g = nkbg.ListGraph.from_edges(10, list(range(9)), list(range(1, 10)), ws)
    data = nkbg.write(g)
    assert (struct.unpack_from("<Q", data, 16)[0] >> 1) & 7 == fmt
    _round_trip(g)


def test_self_loops_both_directions():
    # IOGTest.cpp:953-985
    for directed in (False, True):
        g = nkbg.ListGraph.from_edges(5, range(5), range(5), directed=directed)
        assert _round_trip(g).self_loops() == 5
