"""The threshold random hyperbolic graph generator (GPU) and PageRank parity on it.

Generator: the adjacency built on the GPU from radial bands equals a brute-force all-pairs
test of the same points in numpy (pairs within 1e-12 of the distance threshold excepted), is
symmetric and has no self-loops.  PageRank on an RHG graph (undirected, unweighted, skewed
degrees): FP64 parity with the oracle (same iterations, L1 <= 1e-12) and certified mixed
precision (L1 error <= tol).
"""
import numpy as np
import pytest

import oracle
from networkit_amd import capi
from networkit_amd import generators as gen

pytestmark = pytest.mark.gpu


def brute_force(r, t, R):
    n = len(r)
    adj = []
    ch = np.cosh(R)
    for a in range(n):
        b = np.arange(n)
        lo = np.minimum(a, b)
        hi = np.maximum(a, b)
        s = np.sin(0.5 * (t[lo] - t[hi]))
        val = np.cosh(r[lo] - r[hi]) + 2.0 * np.sinh(r[lo]) * np.sinh(r[hi]) * s * s
        near = np.abs(val / ch - 1.0) < 1e-12
        ok = (val <= ch) & (b != a)
        adj.append((set(np.nonzero(ok & ~near)[0].tolist()), set(np.nonzero(near)[0].tolist())))
    return adj


@pytest.mark.parametrize("n,k,gamma", [(3000, 20.0, 3.0), (2000, 8.0, 2.5)])
def test_rhg_matches_brute_force(n, k, gamma):
    g = gen.rhg(n, k, gamma, seed=3)
    colptr = g["colptr"].cpu().numpy()
    src = g["src"].cpu().numpy()
    r, t, R = gen.rhg_points(n, k, gamma, seed=3)
    ref = brute_force(r, t, R)
    for u in range(n):
        got = set(src[colptr[u]:colptr[u + 1]].tolist())
        sure, near = ref[u]
        assert sure <= got and got <= sure | near, u
        assert u not in got
    # symmetric: every entry appears in both rows
    rows = np.repeat(np.arange(n), np.diff(colptr))
    a = set(zip(rows.tolist(), src.tolist()))
    assert all((v, u) in a for u, v in a)
    # the average degree follows the disk radius (asymptotic formula): same order as asked
    assert 0.3 * k < len(src) / n < 3.0 * k


def test_pagerank_on_rhg():
    n = 100000
    g = gen.rhg(n, 20.0, 3.0, seed=3)
    colptr = g["colptr"].cpu().numpy().astype(np.uint64)
    src = g["src"].cpu().numpy()
    og = oracle.InLists(n, colptr, src.astype(np.uint64), None, directed=False)
    dg = capi.DeviceGraph(n, colptr, src)
    for sinks in (0, 1):
        x = np.full(n, 1.0 / n)
        st = dg.run(x, capi.default_params(sink_handling=sinks))
        ref, it, _ = oracle.pagerank(og, sink_mode=sinks)
        assert st["iterations"] == it
        assert np.abs(x - ref).sum() <= 1e-12
        tight, _, _ = oracle.pagerank(og, tol=1e-15, norm=oracle.L1, max_iterations=20000,
                                      sink_mode=sinks)
        y = np.full(n, 1.0 / n)
        st = dg.run(y, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-8,
                                           stop=capi.NKPR_STOP_L1_ERROR, sink_handling=sinks))
        assert st["l1_error_bound"] <= 1e-8
        assert np.abs(y - tight).sum() <= 1e-8
