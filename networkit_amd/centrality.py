"""networkit.centrality equivalent for the PageRank path.

PageRank, Centrality, Norm (networkit/centrality.pyx:24-100, 1970-2037), EigenvectorCentrality
(:1950-1968) and KatzCentrality with EdgeDirection (:1662-1710) come from the Cython extension.
rankPerNode and relativeRankErrors provide the two pure-Python ranking helpers of
networkit/centrality.pyx:2197-2249 with the same tie rule; unlike those, they accept rankings
whose node ids have gaps (graphs with deleted nodes).
"""
from ._core import (Centrality, EdgeDirection, EigenvectorCentrality,  # noqa: F401
                    KatzCentrality, Norm, PageRank, Precision, SinkHandling)

__all__ = ["Centrality", "EdgeDirection", "EigenvectorCentrality", "KatzCentrality", "Norm",
           "PageRank", "Precision", "SinkHandling", "rankPerNode", "relativeRankErrors"]


def rankPerNode(ranking):
    """Position of every node in ``ranking`` (a list of (node, score) pairs, best first).

    A maximal run of consecutive entries with equal scores shares the position of its first
    entry, the rule of centrality.pyx:2197-2221.  The result is indexed by node id and has
    max(len(ranking), largest id + 1) entries; an id that the ranking does not list (a deleted
    node) holds None."""
    size = len(ranking)
    for node, _ in ranking:
        size = max(size, int(node) + 1)
    position = [None] * size
    run_first = 0
    previous = None
    for i, (node, score) in enumerate(ranking):
        # NaN != NaN, so a NaN score never joins a run (as with the reference's == test)
        if i == 0 or not (score == previous):
            run_first = i
        position[int(node)] = run_first
        previous = score
    return position


def relativeRankErrors(rx, ry):
    """(position in rx + 1) / (position in ry + 1) per node id, from rankPerNode of both
    rankings (centrality.pyx:2223-2249); an empty list when the rankings differ in length.
    Ids that neither ranking lists hold None."""
    if len(rx) != len(ry):
        return []
    px, py = rankPerNode(rx), rankPerNode(ry)
    size = max(len(px), len(py))
    px += [None] * (size - len(px))
    py += [None] * (size - len(py))
    out = []
    for a, b in zip(px, py):
        out.append(None if a is None or b is None else (a + 1) / (b + 1))
    return out
