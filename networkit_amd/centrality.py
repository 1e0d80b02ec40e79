"""networkit.centrality equivalent for the PageRank path.

PageRank, Centrality, Norm (networkit/centrality.pyx:24-100, 1970-2037) come from the Cython
extension; rankPerNode and relativeRankErrors restate the pure-Python helpers of
networkit/centrality.pyx:2197-2249.
"""
from ._core import Centrality, Norm, PageRank, Precision, SinkHandling  # noqa: F401

__all__ = ["Centrality", "Norm", "PageRank", "Precision", "SinkHandling", "rankPerNode",
           "relativeRankErrors"]


def rankPerNode(ranking):
    """For each node id, its position in ``ranking``; nodes with equal scores share the
    position of the first of them (centrality.pyx:2197-2221)."""
    n_nodes = len(ranking)
    ranking_id = [0] * n_nodes
    for index, pair in enumerate(ranking):
        ranking_id[pair[0]] = index
    for index, pair in enumerate(ranking):
        if index == 0:
            continue
        if pair[1] == ranking[index - 1][1]:
            prev_node = ranking[index - 1][0]
            ranking_id[pair[0]] = ranking_id[prev_node]
    return ranking_id


def relativeRankErrors(rx, ry):
    """(r_x(u)+1) / (r_y(u)+1) per node id; empty if the rankings differ in length
    (centrality.pyx:2223-2249)."""
    diff = []
    n = len(rx)
    if not (n == len(ry)):
        return diff
    rnode_x = rankPerNode(rx)
    rnode_y = rankPerNode(ry)
    for i in range(n):
        diff.append((rnode_x[i] + 1) / (rnode_y[i] + 1))
    return diff
