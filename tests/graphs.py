"""Small test graphs shared by the CPU and GPU test modules."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WIKI_VOTE = os.path.join(HERE, "golden", "wiki-Vote.txt.gz")

# The 8-node weighted undirected graph of CentralityGTest.cpp:582-664
#   0    3   6
#    \  / \ /
#     2 -- 5
#    /  \ / \
#   1    4   7
EIGHT_EDGES = [(0, 2, 3.0), (1, 2, 2.0), (2, 3, 3.0), (2, 4, 2.0), (2, 5, 1.5),
               (3, 5, 3.0), (4, 5, 2.0), (5, 6, 3.0), (5, 7, 2.0)]
# Matlab values quoted by the reference test, tolerance 1e-4
EIGHT_PR = [0.0753, 0.0565, 0.2552, 0.1319, 0.0942, 0.2552, 0.0753, 0.0565]
EIGHT_PR_NORMALIZED = [4.0175, 3.0117, 13.6083, 7.0350, 5.0233, 13.6083, 4.0175, 3.0117]

# wiki-Vote known answers (CentralityGTest.cpp:484-539, AlgebraicPageRankGTest.cpp:21-52)
WIKI_TOP = {True: (326, 0.00460), False: (699, 0.00432)}               # +-1e-3
WIKI_TOP_NORMALIZED = {True: (326, 57.6657), False: (699, 205.7324)}    # +-1e-4


def eight_node():
    u = np.array([e[0] for e in EIGHT_EDGES])
    v = np.array([e[1] for e in EIGHT_EDGES])
    w = np.array([e[2] for e in EIGHT_EDGES])
    return 8, u, v, w


def wiki_vote(directed):
    from networkit_amd.graphio import read_snap_edges
    return read_snap_edges(WIKI_VOTE, directed)


def random_graph(n, m, seed, directed=True, weighted=False, self_loops=True, multi=True):
    """A random multigraph edge list (test input only)."""
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    if not self_loops:
        keep = u != v
        u, v = u[keep], v[keep]
    if not multi:
        key = (np.minimum(u, v) * n + np.maximum(u, v)) if not directed else u * n + v
        _, first = np.unique(key, return_index=True)
        first.sort()
        u, v = u[first], v[first]
    w = rng.uniform(0.5, 1.5, len(u)).astype(np.float32).astype(np.float64) if weighted else None
    return n, u, v, w


def skewed_graph(n, m, seed, directed=True, weighted=True, alpha=1.2):
    """Power-law-ish sources and targets: exercises heavy rows and hub columns."""
    rng = np.random.default_rng(seed)
    ranks = np.arange(1, n + 1, dtype=np.float64)
    p = ranks ** (-alpha)
    p /= p.sum()
    perm_a = rng.permutation(n)
    perm_b = rng.permutation(n)
    u = perm_a[rng.choice(n, m, p=p)]
    v = perm_b[rng.choice(n, m, p=p)]
    w = rng.uniform(0.5, 1.5, m).astype(np.float32).astype(np.float64) if weighted else None
    return n, u, v, w
