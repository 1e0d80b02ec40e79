"""ApproximatePageRank and PageRankNibble on the GPU (nkpr_apr_run and the C++/Cython drop-ins)
against the oracle's restatement of the reference (ApproximatePageRank.cpp:18-61,
PageRankNibble.cpp:28-89).

The reference pushes one node at a time from a FIFO queue; the device pushes a frontier at a
time, so the two results are different approximations of the same vector.  What both must
satisfy, checked for each:
  the ACL identity p + pr_alpha(r) = pr_alpha(seeds) (sparse direct solve of the lazy
      personalised PageRank, 1e-12), which fixes the push formula;
  the stopping rule: every residual below eps * weightedDegree(u, true);
  mass conservation (sum p + sum r = 1) and the touched set = support of p and r;
and between the two: ||p_device - p_oracle||_1 <= sum r_device + sum r_oracle (both lie within
their residual mass of the exact vector).  PageRankNibble: the community is the best sweep set
of the device's vector (the oracle's sweep on the same vector gives the same conductance) and
the reference's own checks hold (non-empty, conductance below 1)."""
import numpy as np
import pytest

import oracle
from graphs import random_graph
from networkit_amd import capi
from nkbg_inputs import metis
from test_oracle import lazy_ppr

pytestmark = pytest.mark.gpu


def graph(weighted, seed=9, n=2000, m=12000, self_loops=False):
    z, u, v, w = random_graph(n, m, seed, directed=False, weighted=weighted,
                              self_loops=self_loops)
    ptr, dst, ww = oracle.out_lists(z, u, v, w, directed=False)
    return z, u, v, w, ptr, dst, ww


def check_valid(z, ptr, dst, ww, ids, p, r, seeds, alpha, eps, leaks=False, identity=True,
                edge_pushes=0):
    """identity=False skips the direct solve (its fill-in is too large beyond a few thousand
    nodes of a random graph).  The device rounds every mass it moves to 2^-62 units (at most
    2^-63 off each), which the mass check allows for."""
    P = np.zeros(z)
    R = np.zeros(z)
    P[ids] = p
    R[ids] = r
    vol = oracle.weighted_degree2(z, ptr, dst, ww)
    s = np.zeros(z)
    us = np.unique(seeds)
    s[us] = 1.0 / len(us)
    if identity:
        np.testing.assert_allclose(P + lazy_ppr(z, ptr, dst, ww, R, alpha),
                                   lazy_ppr(z, ptr, dst, ww, s, alpha), rtol=0, atol=1e-12)
    assert (R < eps * vol).all()
    if leaks:
        assert P.sum() + R.sum() < 1.0
    else:
        assert abs(P.sum() + R.sum() - 1.0) <= 1e-12 + edge_pushes * 2.0 ** -63
    assert (P >= 0).all()
    assert np.array_equal(np.sort(ids), ids)
    assert set(ids.tolist()) == set(np.nonzero((P > 0) | (R > 0))[0].tolist())
    return P, R


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("seeds", [[5], [3, 17, 40, 17]])
@pytest.mark.parametrize("alpha,eps", [(0.1, 1e-7), (0.4, 1e-10)])
def test_apr_against_oracle(weighted, seeds, alpha, eps):
    z, u, v, w, ptr, dst, ww = graph(weighted)
    dev = capi.DeviceAPR(z, ptr, dst.astype(np.int32), ww)
    ids, p, r, st = dev.run(seeds, alpha, eps)
    dev.close()
    P, R = check_valid(z, ptr, dst, ww, ids, p, r, seeds, alpha, eps)
    po, ro, _ = oracle.approximate_pagerank(z, ptr, dst, ww, seeds, alpha, eps)
    d = np.abs(P - po).sum()
    print("weighted %s seeds %s alpha %g eps %g: %d rounds, %d pushes, %d touched; "
          "|p - p_oracle|_1 %.2e (bound %.2e)" % (weighted, seeds, alpha, eps, st["rounds"],
                                                  st["pushes"], st["touched"], d,
                                                  R.sum() + ro.sum()), flush=True)
    assert d <= R.sum() + ro.sum() + 1e-12


@pytest.mark.parametrize("weighted", [False, True])
def test_apr_directed_with_sinks(weighted):
    """Directed out-lists with sinks: a sink's residual stays (no neighbour's add queues it), a
    sink seed is pushed out (losing what it cannot send); the device vector satisfies the
    identity for the walk whose sink rows are the lazy half alone, loses no other mass, and lies
    within the two runs' residual mass of the oracle's."""
    z, u, v, w = random_graph(3000, 9000, 19, directed=True, weighted=weighted, self_loops=False)
    ptr, dst, ww = oracle.out_lists(z, u, v, w, directed=True)
    vol = oracle.weighted_degree2(z, ptr, dst, ww)
    sinks = np.nonzero(vol == 0)[0]
    assert len(sinks) > 50
    seeds = [int(np.nonzero(vol > 0)[0][3]), int(sinks[2])]
    alpha, eps = 0.15, 1e-9
    dev = capi.DeviceAPR(z, ptr, dst.astype(np.int32), ww)
    ids, p, r, st = dev.run(seeds, alpha, eps)
    dev.close()
    P = np.zeros(z)
    R = np.zeros(z)
    P[ids] = p
    R[ids] = r
    s = np.zeros(z)
    s[np.unique(seeds)] = 0.5
    np.testing.assert_allclose(P + lazy_ppr(z, ptr, dst, ww, R, alpha),
                               lazy_ppr(z, ptr, dst, ww, s, alpha), rtol=0, atol=1e-12)
    # every push of the sink seed adds alpha res to p and has nowhere to send (1 - alpha) res / 2:
    # (1 - alpha) / (2 alpha) p[seed] is lost in all; every other push conserves mass
    lost = (1 - alpha) / (2 * alpha) * P[seeds[1]]
    assert lost > 0.3 and abs(P.sum() + R.sum() - (1.0 - lost)) <= 1e-12
    assert (R[vol > 0] < eps * vol[vol > 0]).all()
    assert R[seeds[1]] == 0.0 and R[sinks].sum() > 0
    po, ro, _ = oracle.approximate_pagerank(z, ptr, dst, ww, seeds, alpha, eps)
    d = np.abs(P - po).sum()
    print("directed, weighted %s: %d rounds, %d touched, |p - p_oracle|_1 %.2e (bound %.2e)" % (
        weighted, st["rounds"], st["touched"], d, R.sum() + ro.sum()), flush=True)
    assert d <= R.sum() + ro.sum() + 1e-12


@pytest.mark.parametrize("weighted", [False, True])
def test_apr_self_loops(weighted):
    """Self-loops: a push sends nothing through one (the reference overwrites r[u] after its
    loop), so the identity holds for the walk without them and mass leaks, on the device as in
    the oracle; the two vectors stay within their residual and leaked mass of each other."""
    z, u, v, w, ptr, dst, ww = graph(weighted, seed=12, n=1500, m=9000, self_loops=True)
    loops = int((np.repeat(np.arange(z), np.diff(ptr.astype(np.int64))) == dst).sum())
    assert loops > 0
    seeds, alpha, eps = [4, 9], 0.15, 1e-9
    dev = capi.DeviceAPR(z, ptr, dst.astype(np.int32), ww)
    ids, p, r, st = dev.run(seeds, alpha, eps)
    dev.close()
    P, R = check_valid(z, ptr, dst, ww, ids, p, r, seeds, alpha, eps, leaks=True)
    po, ro, _ = oracle.approximate_pagerank(z, ptr, dst, ww, seeds, alpha, eps)
    # p_x = pr(s) - pr(r_x) for both, and pr is a contraction in L1 on non-negative vectors
    d = np.abs(P - po).sum()
    print("self-loops %d, weighted %s: mass %.6f (oracle %.6f), |p - p_oracle|_1 %.2e" % (
        loops, weighted, P.sum() + R.sum(), po.sum() + ro.sum(), d), flush=True)
    assert d <= R.sum() + ro.sum() + 1e-12


def test_apr_reference_inputs_drop_in():
    """test_scd.py:12-25 through the Cython class: scores of existing nodes, non-negative, the
    same for seed 0 and [0]; and SelectiveCDGTest.cpp:33-39 (wiki-Vote is exercised by the
    other tests; here hep-th with seed 50)."""
    import networkit_amd as nk
    for name, seed, alpha, eps in (("PGPgiantcompo.graph", 0, 0.1, 1e-12),
                                   ("hep-th.graph", 50, 0.4, 1e-12)):
        n, u, v = metis(name)
        G = nk.Graph(n, False, False)
        G.addEdges(u, v)
        a = nk.scd.ApproximatePageRank(G, alpha, eps)
        r1 = dict(a.run(seed))
        r2 = dict(nk.scd.ApproximatePageRank(G, alpha, eps).run([seed]))
        assert len(r1) >= 1 and len(r1) == len(r2)
        for x, sc in r1.items():
            assert G.hasNode(x) and sc >= 0 and x in r2
            assert sc == r2[x]                 # residual sums are exact: the same bits
        assert r1[seed] > 0


def test_pagerank_nibble():
    """The community is the best sweep set of the pushes' vector (the oracle's sweep on a
    device vector gives the same best conductance) and passes SelectiveCDGTest.cpp:70-125's
    checks: non-empty, conductance below 1."""
    import networkit_amd as nk
    n, u, v = metis("hep-th.graph")
    G = nk.Graph(n, False, False)
    G.addEdges(u, v)
    ptr, dst, ww = oracle.out_lists(n, u, v, None, directed=False)
    alpha, eps = 0.1, 1e-5
    com = nk.scd.PageRankNibble(G, alpha, eps).run([50])[50]
    assert len(com) > 0
    vol = oracle.weighted_degree2(n, ptr, dst, ww)
    total = 2.0 * len(u)

    def conductance(S):
        pi = np.asarray(ptr, np.int64)
        cut = volume = 0.0
        for x in S:
            for e in range(pi[x], pi[x + 1]):
                volume += 1.0
                if int(dst[e]) not in S or int(dst[e]) == x:
                    cut += 1.0
        return cut / min(volume, total - volume)

    c = conductance(com)
    assert c < 1.0
    dev = capi.DeviceAPR(n, ptr, dst.astype(np.int32), None)
    ids, p, r, _ = dev.run([50], alpha, eps)
    dev.close()
    order, best, cond = oracle.best_sweep_set(n, ptr, dst, ww, n, total, ids.astype(np.uint64),
                                              p / vol[ids])
    assert abs(cond - c) <= 1e-9 * c
    print("PageRankNibble: community of %d nodes, conductance %.4f" % (len(com), c), flush=True)


def test_apr_errors():
    z, u, v, w, ptr, dst, ww = graph(False, n=100, m=400)
    dev = capi.DeviceAPR(z, ptr, dst.astype(np.int32), None)
    with pytest.raises(capi.NkprError, match="seed"):
        dev.run([z + 5], 0.1, 1e-6)
    with pytest.raises(capi.NkprError, match="alpha"):
        dev.run([0], 1.5, 1e-6)
    dev.close()
    bad_w = np.ones(len(dst))
    bad_w[7] = -1.0
    with pytest.raises(capi.NkprError, match="non-negative"):
        capi.DeviceAPR(z, ptr, dst.astype(np.int32), bad_w)
    bad_dst = dst.astype(np.int32).copy()
    bad_dst[3] = z
    with pytest.raises(capi.NkprError, match="ids"):
        capi.DeviceAPR(z, ptr, bad_dst, None)
    bad_ptr = ptr.copy()
    bad_ptr[10], bad_ptr[11] = bad_ptr[11], bad_ptr[10] + 1
    with pytest.raises(capi.NkprError, match="offsets"):
        capi.DeviceAPR(z, bad_ptr, dst.astype(np.int32), None)


def test_apr_runs_are_reproducible_and_independent():
    """The same run gives the same bits every time, also after runs from other seeds on the
    same handle (only the entries a run touched are cleared afterwards); a node without
    neighbours as the seed is pushed until its residual is gone, as in the reference (the
    oracle's vol-0 rule), and other seeds of the same run are unaffected by it."""
    z, u, v, w, ptr, dst, ww = graph(True, seed=21, n=60000, m=600000)
    dev = capi.DeviceAPR(z, ptr, dst.astype(np.int32), ww)
    a = dev.run([11, 500], 0.05, 1e-10)
    dev.run([7], 0.3, 1e-8)
    dev.run(list(range(0, 60000, 7)), 0.1, 1e-7)
    b = dev.run([11, 500], 0.05, 1e-10)
    dev.close()
    print("%d rounds, %d pushes, %d touched" % (a[3]["rounds"], a[3]["pushes"], a[3]["touched"]),
          flush=True)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    assert a[3]["pushes"] == b[3]["pushes"] and a[3]["rounds"] == b[3]["rounds"]
    check_valid(z, ptr, dst, ww, a[0], a[1], a[2], [11, 500], 0.05, 1e-10, identity=False,
                edge_pushes=a[3]["edge_pushes"])
    # an isolated node (id z) next to a connected seed
    ptr2 = np.concatenate([ptr, ptr[-1:]])
    dev = capi.DeviceAPR(z + 1, ptr2, dst.astype(np.int32), ww)
    ids, p, r, st = dev.run([z, 11], 0.2, 1e-9)
    dev.close()
    po, ro, _ = oracle.approximate_pagerank(z + 1, ptr2, dst, ww, [z, 11], 0.2, 1e-9)
    at = int(np.searchsorted(ids, z))
    assert ids[at] == z and r[at] == 0.0
    assert abs(p[at] - po[z]) <= 1e-15 and abs(po[z] - 0.5 * 0.2 / (1 - 0.4)) <= 1e-15
