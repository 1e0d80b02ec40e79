"""PageRank of graphs read from NetworKit binary files: nkpr_nkbg_read's pull lists go to the
device as they are (nkpr_graph_create) and the run matches the oracle on the lists the
reference's reader builds (oracle/nkbg.py): the oracle's iteration count and L1 <= 1e-12."""
import numpy as np
import pytest

import oracle
from graphs import random_graph
from networkit_amd import capi
from nkbg_inputs import konect, wiki_vote_edges
from oracle import nkbg

pytestmark = pytest.mark.gpu


def cases():
    z, u, v = wiki_vote_edges()
    yield "wiki_directed", nkbg.ListGraph.from_edges(z, u, v, directed=True)
    n, u, v, w = konect("foodweb-baydry.konect")
    yield "konect_weighted", nkbg.ListGraph.from_edges(n, u, v, w, directed=True)
    z, u, v, w = random_graph(5000, 40000, 8, directed=False, weighted=True)
    yield "undirected_weighted_loops", nkbg.ListGraph.from_edges(z, u, v, w)


@pytest.mark.parametrize("name", ["wiki_directed", "konect_weighted",
                                  "undirected_weighted_loops"])
def test_pagerank_from_nkbg_file(tmp_path, name):
    g = dict(cases())[name]
    path = tmp_path / "g.nkbg"
    path.write_bytes(nkbg.write(g, chunks=13))
    back, _, _ = nkbg.read(path.read_bytes())
    lists = back.pull_lists()
    ptr = np.zeros(back.n + 1, np.uint64)
    ptr[1:] = np.cumsum([len(x) for x in lists])
    src = np.array([v for x in lists for v, _ in x], np.uint64)
    w = np.array([c for x in lists for _, c in x]) if back.weighted else None
    ref, it, _ = oracle.pagerank(oracle.InLists(back.n, ptr, src, w))
    f = capi.read_nkbg(path)
    dg = capi.DeviceGraph(f["nodes"], f["colptr"], f["src"],
                          None if f["weight"] is None else f["weight"].astype(np.float32),
                          weight64=f["weight"], exists=f["exists"])
    x = np.full(f["nodes"], 1.0 / f["nodes"])
    st = dg.run(x)
    dg.close()
    print("%s: %d iterations (oracle %d), L1 %.3e" % (name, st["iterations"], it,
                                                      np.abs(x - ref).sum()), flush=True)
    assert st["iterations"] == it
    assert np.abs(x - ref).sum() <= 1e-12
