"""NetworKit-style API semantics that need no GPU: Graph storage rules, the PageRank
properties, errors before run(), and the pure-Python rank helpers."""
import pytest

import networkit_amd as nk
from graphs import EIGHT_EDGES


def eight():
    G = nk.Graph(8, True, False)
    for u, v, w in EIGHT_EDGES:
        G.addEdge(u, v, w)
    return G


def test_graph_counts_and_degrees():
    G = eight()
    assert G.numberOfNodes() == 8 and G.numberOfEdges() == 9 and G.upperNodeIdBound() == 8
    assert G.isWeighted() and not G.isDirected()
    assert G.degree(2) == 5 and G.weightedDegree(2) == 3 + 2 + 3 + 2 + 1.5
    assert G.hasEdge(2, 0) and G.hasEdge(0, 2) and not G.hasEdge(0, 1)


def test_self_loop_stored_once_undirected():
    G = nk.Graph(2, True, False)
    G.addEdge(0, 0, 2.5)
    G.addEdge(0, 1, 1.0)
    assert G.degree(0) == 2 and G.weightedDegree(0) == 3.5
    assert G.weightedDegree(0, True) == 6.0
    assert G.numberOfSelfLoops() == 1


def test_directed_degrees_and_multi_edges():
    G = nk.Graph(3, False, True)
    G.addEdges([0, 0, 1], [1, 1, 2])
    assert G.numberOfEdges() == 3
    assert G.degreeOut(0) == 2 and G.degreeIn(1) == 2 and G.degreeIn(0) == 0
    assert G.weightedDegree(0) == 2.0


def test_remove_and_restore_node():
    G = nk.Graph(4, False, True)
    G.addEdges([0, 1, 2], [1, 2, 3])
    G.removeNode(2)
    assert not G.hasNode(2) and G.numberOfNodes() == 3 and G.numberOfEdges() == 1
    G.restoreNode(2)
    assert G.hasNode(2) and G.degree(2) == 0
    with pytest.raises(RuntimeError):
        G.restoreNode(2)


def test_pagerank_properties_before_run():
    G = eight()
    pr = nk.centrality.PageRank(G)
    assert pr.norm == nk.centrality.Norm.l2norm
    assert pr.maxIterations == 2 ** 64 - 1
    pr.norm = nk.centrality.Norm.l1norm
    assert pr.norm == 0
    pr.maxIterations = 7
    assert pr.maxIterations == 7
    with pytest.raises(Exception):
        pr.maxIterations = -1
    assert not pr.hasFinished()
    for f in (pr.scores, pr.ranking, pr.numberOfIterations, pr.centralization):
        with pytest.raises(RuntimeError):
            f()
    with pytest.raises(RuntimeError):
        pr.score(0)
    assert pr.precision == nk.centrality.Precision.FP64


def test_abstract_bases_refuse_direct_use():
    with pytest.raises(RuntimeError):
        nk.base.Algorithm()
    with pytest.raises(RuntimeError):
        nk.centrality.Centrality()


def test_rank_helpers():
    r = [(2, 0.5), (0, 0.3), (1, 0.3), (3, 0.1)]
    assert nk.centrality.rankPerNode(r) == [1, 1, 0, 3]
    assert nk.centrality.relativeRankErrors(r, r) == [1.0, 1.0, 1.0, 1.0]
    assert nk.centrality.relativeRankErrors(r, r[:2]) == []


def _sorted_ranking(rng, ids, levels):
    scores = rng.integers(0, levels, len(ids)) / 8.0
    return sorted(zip(ids.tolist(), scores.tolist()), key=lambda p: (-p[1], p[0]))


def test_rank_helpers_brute_force():
    """On a ranking sorted by score, the shared position of a run of equal scores is the number
    of entries with a strictly higher score; deleted ids (gaps) hold None."""
    import numpy as np
    rng = np.random.default_rng(11)
    for trial in range(40):
        z = int(rng.integers(1, 60))
        keep = rng.random(z) < (1.0 if trial % 2 else 0.7)
        keep[rng.integers(0, z)] = True
        ids = np.nonzero(keep)[0]
        r = _sorted_ranking(rng, ids, int(rng.integers(1, 6)))
        pos = nk.centrality.rankPerNode(r)
        assert len(pos) == max(len(r), int(ids.max()) + 1)
        score = dict(r)
        for u in range(len(pos)):
            if u in score:
                assert pos[u] == sum(1 for _, s in r if s > score[u])
            else:
                assert pos[u] is None
        r2 = _sorted_ranking(rng, ids, 3)
        rel = nk.centrality.relativeRankErrors(r, r2)
        pos2 = nk.centrality.rankPerNode(r2)
        for u in range(len(rel)):
            if u in score:
                assert rel[u] == (pos[u] + 1) / (pos2[u] + 1)
            else:
                assert rel[u] is None


def test_graph_accessors_reject_unknown_ids():
    G = eight()
    for f in (G.degree, G.degreeIn, G.degreeOut, G.weightedDegree):
        with pytest.raises(IndexError):
            f(8)
    with pytest.raises(IndexError):
        G.degree(2 ** 40)
    assert not G.hasEdge(0, 2 ** 40)
    assert G.weight(0, 2 ** 40) == 0.0
    with pytest.raises(RuntimeError):
        G.addEdge(0, 8)


def test_snap_reader_ids_and_dedup(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# c\n10 20\n20 10\n10 20\n30\t10\n")
    from networkit_amd.graphio import read_snap_edges
    z, u, v = read_snap_edges(str(p), directed=True)
    assert z == 3 and u.tolist() == [0, 1, 2] and v.tolist() == [1, 0, 0]
    z, u, v = read_snap_edges(str(p), directed=False)
    assert u.tolist() == [0, 2] and v.tolist() == [1, 0]


def test_bench_refuses_more_than_one_gpu():
    """The benchmark measures one GPU: a request for more exits with status 2 and says so,
    before anything touches a device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       cwd=root, capture_output=True, text=True, timeout=300)
    assert p.returncode == 2
    assert "measures one GPU" in p.stderr and p.stdout == ""


def test_eigenvector_and_katz_before_run():
    G = eight()
    ev = nk.centrality.EigenvectorCentrality(G)
    kc = nk.centrality.KatzCentrality(G)
    for c in (ev, kc):
        assert not c.hasFinished()
        with pytest.raises(RuntimeError):
            c.scores()
        with pytest.raises(RuntimeError):
            c.ranking()
    assert kc.edgeDirection == nk.centrality.EdgeDirection.inEdges == 0
    kc.edgeDirection = nk.centrality.EdgeDirection.outEdges
    assert kc.edgeDirection == 1
    with pytest.raises(ValueError):
        kc.edgeDirection = 5
    with pytest.raises(RuntimeError, match="Alpha must be non negative"):
        nk.centrality.KatzCentrality(G, -0.5)
