"""Readers of the reference's small text fixtures (test input only) for the nkbg tests:
METIS (tests/golden/tiny_01.graph, jazz.graph: undirected, 1-based neighbour lists, each edge
listed at both endpoints) and KONECT (foodweb-baydry.konect: directed, weighted, 1-based)."""
import gzip
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def metis(name):
    """(n, u, v): every edge once, as (smaller id, larger id) in line order."""
    with open(os.path.join(GOLDEN, name)) as f:
        lines = [ln for ln in f if not ln.lstrip().startswith("%")]
    n = int(lines[0].split()[0])
    u, v = [], []
    for a, ln in enumerate(lines[1:n + 1]):
        for t in ln.split():
            b = int(t) - 1
            if a <= b:
                u.append(a)
                v.append(b)
    return n, np.array(u, np.int64), np.array(v, np.int64)


def konect(name):
    """(n, u, v, w) of a directed weighted KONECT file (ids 1-based)."""
    u, v, w = [], [], []
    with open(os.path.join(GOLDEN, name)) as f:
        for ln in f:
            if ln.startswith("%") or not ln.strip():
                continue
            t = ln.split()
            u.append(int(t[0]) - 1)
            v.append(int(t[1]) - 1)
            w.append(float(t[2]))
    n = max(max(u), max(v)) + 1
    return n, np.array(u, np.int64), np.array(v, np.int64), np.array(w)


def wiki_vote_edges():
    from networkit_amd.graphio import read_snap_edges
    return read_snap_edges(os.path.join(GOLDEN, "wiki-Vote.txt.gz"), True)
