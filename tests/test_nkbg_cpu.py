"""The library's nkbg reader and writer (nkpr_nkbg_read / nkpr_nkbg_write and the C++ drop-ins
NetworkitBinaryReader / NetworkitBinaryWriter) against the nkbg oracle (oracle/nkbg.py).
Host code only: these run without a GPU."""
import os

import numpy as np
import pytest

import networkit_amd as nk
from graphs import random_graph
from networkit_amd import capi
from nkbg_inputs import konect, metis, wiki_vote_edges
from oracle import nkbg


def cases():
    n, u, v = metis("jazz.graph")
    yield "jazz", nkbg.ListGraph.from_edges(n, u, v)
    n, u, v, w = konect("foodweb-baydry.konect")
    yield "konect", nkbg.ListGraph.from_edges(n, u, v, w, directed=True)
    z, u, v = wiki_vote_edges()
    yield "wiki", nkbg.ListGraph.from_edges(z, u, v, directed=True)
    z, u, v, w = random_graph(300, 3000, 4, directed=False, weighted=True)
    yield "undirected_float_loops", nkbg.ListGraph.from_edges(z, u, v, w)
    z, u, v, _ = random_graph(200, 2500, 5, directed=True)
    yield "directed_signed", nkbg.ListGraph.from_edges(
        z, u, v, np.random.default_rng(5).integers(-9, 9, len(u)).astype(float), directed=True)
    z, u, v, _ = random_graph(100, 800, 6, directed=False)
    yield "varint_weights", nkbg.ListGraph.from_edges(
        z, u, v, np.random.default_rng(6).integers(0, 1 << 50, len(u)).astype(float))
    z, u, v, _ = random_graph(100, 800, 7, directed=True)
    yield "double_weights", nkbg.ListGraph.from_edges(
        z, u, v, np.random.default_rng(7).random(len(u)), directed=True)


CASES = {k: g for k, g in cases()}


def lists_of(g):
    """(ptr, ids, weights) of the pull lists and, when directed, the out-lists."""
    def flat(lists):
        ptr = np.zeros(g.n + 1, np.uint64)
        ptr[1:] = np.cumsum([len(x) for x in lists])
        ids = np.array([v for x in lists for v, _ in x], np.int64)
        ws = np.array([w for x in lists for _, w in x], np.float64)
        return ptr, ids, ws
    return flat(g.pull_lists()), (flat(g.out) if g.directed else None)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("chunks", [32, 1, 7])
@pytest.mark.parametrize("with_out", [True, False])
def test_library_reader_matches_oracle(tmp_path, name, chunks, with_out):
    """Pull lists, out-lists, edge and self-loop counts equal the oracle reader's; without
    out-lists a directed file's out-lists are stepped over and its self-loops counted in the
    in-lists."""
    g = CASES[name]
    path = tmp_path / "g.nkbg"
    path.write_bytes(nkbg.write(g, chunks=chunks))
    back, edges, loops = nkbg.read(path.read_bytes())
    got = capi.read_nkbg(path, with_out_lists=with_out)
    assert got["nodes"] == g.n and got["edges"] == edges and got["self_loops"] == loops
    assert got["directed"] == g.directed and got["weighted"] == g.weighted
    (ptr, ids, ws), outl = lists_of(back)
    assert np.array_equal(got["colptr"], ptr)
    assert np.array_equal(got["src"], ids)
    if g.weighted:
        assert np.array_equal(got["weight"], ws)
    else:
        assert got["weight"] is None
    if g.directed and with_out:
        assert np.array_equal(got["out_ptr"], outl[0])
        assert np.array_equal(got["out_dst"], outl[1])
    assert got["exists"].all()


@pytest.mark.parametrize("name", list(CASES))
def test_library_writer_bytes_match_oracle(tmp_path, name):
    g = CASES[name]
    (ptr, ids, ws), _ = lists_of(g)
    out = [(v, w) for lst in g.out for v, w in lst]
    optr = np.zeros(g.n + 1, np.uint64)
    optr[1:] = np.cumsum([len(x) for x in g.out])
    path = tmp_path / "w.nkbg"
    capi.write_nkbg(path, g.n, g.directed, optr, [v for v, _ in out],
                    [w for _, w in out] if g.weighted else None,
                    ptr if g.directed else None, ids if g.directed else None,
                    ws if (g.directed and g.weighted) else None)
    assert path.read_bytes() == nkbg.write(g)


def test_drop_in_reader_and_writer_round_trip(tmp_path):
    """IOGTest.cpp:822-905 on the C++ classes: same node and edge counts, every edge present
    with its weight, self-loops kept."""
    for name in ("konect", "undirected_float_loops", "wiki"):
        g = CASES[name]
        path = tmp_path / (name + ".nkbg")
        path.write_bytes(nkbg.write(g))
        G = nk.graphio.NetworkitBinaryReader().read(path)
        assert G.numberOfNodes() == g.n and G.isDirected() == g.directed
        assert G.isWeighted() == g.weighted
        assert G.numberOfEdges() == g.number_of_edges()
        assert G.numberOfSelfLoops() == g.self_loops()
        for u in range(g.n):
            for v, w in g.out[u][:3]:
                assert G.hasEdge(u, v)
                if g.weighted:
                    assert G.weight(u, v) == next(x for y, x in g.out[u] if y == v)
        # writing the Graph again gives the file the reference's writer gives for it
        path2 = tmp_path / (name + "2.nkbg")
        nk.graphio.NetworkitBinaryWriter().write(G, path2)
        back, _, _ = nkbg.read(path2.read_bytes())
        assert nkbg.write(back) == path2.read_bytes()


def test_deleted_slots_and_errors(tmp_path):
    G = nk.Graph(6, True, True)
    G.addEdges([0, 1, 2, 5], [1, 2, 5, 0], [1.5, 2.5, 3.5, 4.5])
    G.removeNode(3)
    path = tmp_path / "d.nkbg"
    nk.graphio.NetworkitBinaryWriter().write(G, path)
    got = capi.read_nkbg(path)
    assert got["exists"].tolist() == [1, 1, 1, 0, 1, 1]
    G2 = nk.graphio.NetworkitBinaryReader().read(path)
    assert G2.numberOfNodes() == 5 and not G2.hasNode(3) and G2.hasEdge(2, 5)
    data = path.read_bytes()
    for bad in (b"nkbg001\0" + data[8:], data[:60], data[:len(data) - 3]):
        p = tmp_path / "bad.nkbg"
        p.write_bytes(bad)
        with pytest.raises(capi.NkprError):
            capi.read_nkbg(p)
        with pytest.raises(RuntimeError):
            nk.graphio.NetworkitBinaryReader().read(p)
    with pytest.raises(capi.NkprError):
        capi.read_nkbg(tmp_path / "missing.nkbg")


def test_empty_file_is_an_error(tmp_path):
    path = tmp_path / "e.nkbg"
    path.write_bytes(b"")
    with pytest.raises(capi.NkprError, match="magic"):
        capi.read_nkbg(path)
