"""Pins the CPU oracle (oracle/pagerank_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement fails it:
  - reference known answers (8-node Matlab values, wiki-Vote top-1 in four variants) catch a
    wrong in-edge direction, a missing damping/teleport term or a wrong normalisation;
  - the sparse direct solve catches a dropped or transposed term at 1e-9 instead of 1e-4;
  - dense power iteration pins the sink-distribution mode, which the reference lacks;
  - invariants (uniform scores on cycles and regular graphs, automorphic pairs, iteration cap,
    warm start) pin the loop control.
"""
import numpy as np
import pytest

import oracle
from closed_forms import dense_power_iteration, pagerank_closed_form, solve_fixed_point
from graphs import (EIGHT_PR, EIGHT_PR_NORMALIZED, WIKI_TOP, WIKI_TOP_NORMALIZED, eight_node,
                    random_graph, skewed_graph, wiki_vote)


def test_in_lists_directed_order():
    g = oracle.in_lists(3, [0, 1, 0, 2], [1, 2, 2, 2], directed=True)
    assert g.in_ptr.tolist() == [0, 0, 1, 4]
    assert g.in_src.tolist() == [0, 1, 0, 2]


def test_in_lists_undirected_selfloop_once():
    g = oracle.in_lists(3, [0, 1, 1], [1, 1, 2], [2.0, 5.0, 7.0], directed=False)
    # node 0: {1}; node 1: {0, 1 (loop once), 2}; node 2: {1}
    assert g.in_ptr.tolist() == [0, 1, 4, 5]
    assert g.in_src.tolist() == [1, 0, 1, 2, 1]
    assert g.in_w.tolist() == [2.0, 2.0, 5.0, 7.0, 7.0]
    deg, cnt = oracle.out_degrees(g)
    assert deg.tolist() == [2.0, 14.0, 7.0]
    assert cnt.tolist() == [1, 3, 1]


def test_eight_node_known_answers():
    z, u, v, w = eight_node()
    g = oracle.in_lists(z, u, v, w, directed=False)
    s, it, mx = oracle.pagerank(g, 0.85)
    assert np.allclose(s, EIGHT_PR, atol=1e-4)
    assert mx == 1.0
    s, it, mx = oracle.pagerank(g, 0.85, normalized=True)
    assert np.allclose(s, EIGHT_PR_NORMALIZED, atol=1e-4)
    assert mx == pytest.approx(s.max())


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("norm", [oracle.L1, oracle.L2])
def test_wiki_vote_known_answers(directed, norm):
    z, u, v = wiki_vote(directed)
    g = oracle.in_lists(z, u, v, None, directed)
    s, it, _ = oracle.pagerank(g, norm=norm)
    r = oracle.ranking(s)
    top, val = WIKI_TOP[directed]
    assert r[0] == top and abs(s[top] - val) <= 1e-3
    s2, it2, _ = oracle.pagerank(g, norm=norm, max_iterations=2, start=s)
    assert it2 <= 2
    sn, _, mx = oracle.pagerank(g, norm=norm, normalized=True)
    rn = oracle.ranking(sn)
    topn, valn = WIKI_TOP_NORMALIZED[directed]
    assert rn[0] == topn and abs(sn[topn] - valn) <= 1e-4
    assert mx == sn.max()


@pytest.mark.parametrize("case", ["er_dir", "er_undir_w", "skew_dir_w", "wiki_dir"])
@pytest.mark.parametrize("sink_mode", [0, 1])
def test_oracle_matches_direct_solve(case, sink_mode):
    if case == "er_dir":
        z, u, v, w = random_graph(3000, 9000, 1, directed=True)
        g = oracle.in_lists(z, u, v, w, True)
    elif case == "er_undir_w":
        z, u, v, w = random_graph(2000, 6000, 2, directed=False, weighted=True)
        g = oracle.in_lists(z, u, v, w, False)
    elif case == "skew_dir_w":
        z, u, v, w = skewed_graph(4000, 30000, 3)
        g = oracle.in_lists(z, u, v, w, True)
    else:
        z, u, v = wiki_vote(True)
        g = oracle.in_lists(z, u, v, None, True)
    s, it, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=5000,
                               sink_mode=sink_mode)
    ref = pagerank_closed_form(g, 0.85, sink_mode)
    assert np.abs(s - ref).sum() < 1e-11
    # a wrong edge direction (transposed operator) is caught
    gt = oracle.in_lists(z, g_src_rows(g)[1], g_src_rows(g)[0], g.in_w, True)
    st, _, _ = oracle.pagerank(gt, tol=1e-15, norm=oracle.L1, max_iterations=5000,
                               sink_mode=sink_mode)
    if case != "er_undir_w":
        assert np.abs(st - ref).sum() > 1e-3


def g_src_rows(g):
    rows = np.repeat(np.arange(g.z), np.diff(g.in_ptr.astype(np.int64)))
    return g.in_src.astype(np.int64), rows


@pytest.mark.parametrize("sink_mode", [0, 1])
def test_oracle_matches_dense_power(sink_mode):
    z, u, v, w = random_graph(60, 150, 7, directed=True, weighted=True)
    g = oracle.in_lists(z, u, v, w, True)
    s, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=10000,
                              sink_mode=sink_mode)
    ref = dense_power_iteration(g, 0.85, sink_mode)
    assert np.abs(s - ref).sum() < 1e-12


def test_distribute_sinks_keeps_mass():
    z, u, v, w = random_graph(500, 700, 11, directed=True)
    g = oracle.in_lists(z, u, v, w, True)
    deg, cnt = oracle.out_degrees(g)
    assert (cnt == 0).sum() > 10
    s, _, _ = oracle.pagerank(g, tol=1e-14, norm=oracle.L1, max_iterations=1,
                              sink_mode=1)
    # one sweep from 1/n keeps total mass 1 exactly (up to rounding) before normalisation
    x0 = np.full(z, 1.0 / z)
    sinks = cnt == 0
    import closed_forms
    P, idx, exists, snk, n = closed_forms._operator(g)
    y = 0.85 * (P @ x0 + x0[sinks].sum() / n) + 0.15 / n
    assert abs(y.sum() - 1.0) < 1e-12
    assert np.allclose(s, y / y.sum(), rtol=1e-12, atol=0)


def test_uniform_on_cycle_and_regular():
    n = 97
    u = np.arange(n)
    g = oracle.in_lists(n, u, (u + 1) % n, None, True)
    s, _, _ = oracle.pagerank(g)
    assert np.allclose(s, 1.0 / n, rtol=0, atol=1e-15)
    g = oracle.in_lists(n, np.concatenate([u, u]), np.concatenate([(u + 1) % n, (u + 5) % n]),
                        None, False)
    s, _, _ = oracle.pagerank(g)
    assert np.allclose(s, 1.0 / n, rtol=0, atol=1e-15)


def test_automorphic_pairs_equal():
    z, u, v, w = eight_node()
    g = oracle.in_lists(z, u, v, w, directed=False)
    s, _, _ = oracle.pagerank(g, tol=1e-14)
    for a, b in [(0, 6), (1, 7), (2, 5)]:
        assert s[a] == pytest.approx(s[b], rel=1e-12)


def test_no_in_edges_gets_teleport():
    # node 0 has no in-edges: after any sweep its unnormalised score is (1-d)/n exactly
    g = oracle.in_lists(4, [0, 1, 2, 3], [1, 2, 3, 1], None, True)
    s, it, _ = oracle.pagerank(g, max_iterations=1)
    # single sweep from uniform: node 0 -> 0.15/4, normalised by the sum
    y = np.array([0.0375, 0.85 * (0.25 + 0.25) + 0.0375, 0.85 * 0.25 + 0.0375,
                  0.85 * 0.25 + 0.0375])
    assert np.allclose(s, y / y.sum(), rtol=1e-15)
    assert it == 1


def test_max_iterations_zero_still_sweeps_once():
    z, u, v, w = random_graph(50, 100, 3)
    g = oracle.in_lists(z, u, v, w, True)
    _, it, _ = oracle.pagerank(g, max_iterations=0)
    assert it == 1


def test_deleted_slots_keep_start_value():
    z, u, v, w = random_graph(40, 120, 4, directed=True)
    keep = (u != 7) & (v != 7) & (u != 13) & (v != 13)
    exists = np.ones(z, np.uint8)
    exists[[7, 13]] = 0
    g = oracle.in_lists(z, u[keep], v[keep], None, True, exists=exists)
    s, _, mx = oracle.pagerank(g, normalized=True, tol=1e-15, norm=oracle.L1,
                               max_iterations=5000)
    assert s[7] == 1.0 / 38 and s[13] == 1.0 / 38
    ref = pagerank_closed_form(g, 0.85, 0, normalized=True)
    assert np.abs(s[exists == 1] - ref[exists == 1]).sum() < 1e-11


def test_ranking_bruteforce():
    rng = np.random.default_rng(5)
    s = rng.integers(0, 20, 500).astype(np.float64) / 7.0
    exists = (rng.random(500) > 0.1).astype(np.uint8)
    r = oracle.ranking(s, exists)
    ids = [i for i in range(500) if exists[i]]
    ids.sort(key=lambda i: (-s[i], i))
    assert r.tolist() == ids


# ---- personalised teleport (batched PPR oracle) ------------------------------------------------

def _teleport(z, seeds, weights=None, exists=None):
    t = np.zeros(z)
    np.add.at(t, np.asarray(seeds), 1.0 if weights is None else np.asarray(weights, float))
    return t / t.sum()


@pytest.mark.parametrize("sink_mode", [0, 1])
def test_oracle_teleport_matches_direct_solve(sink_mode):
    z, u, v, w = skewed_graph(3000, 20000, 21)
    g = oracle.in_lists(z, u, v, w, True)
    rng = np.random.default_rng(3)
    for k in range(4):
        seeds = rng.choice(z, rng.integers(1, 40), replace=False)
        t = _teleport(z, seeds, rng.uniform(0.1, 2.0, len(seeds)) if k % 2 else None)
        s, it, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=5000,
                                   sink_mode=sink_mode, teleport=t)
        x, exists, sinks, n = solve_fixed_point(g, 0.85, sink_mode, teleport=0.15 * t)
        ref = x / x.sum()
        assert np.abs(s - ref).sum() < 1e-11
        # the teleport enters as (1-d) t: a run with the uniform rule instead differs
        su, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=5000,
                                   sink_mode=sink_mode)
        assert np.abs(su - ref).sum() > 1e-2


def test_oracle_uniform_teleport_is_reference_loop():
    z, u, v, w = random_graph(2000, 7000, 5, directed=True, weighted=True)
    g = oracle.in_lists(z, u, v, w, True)
    for sink_mode in (0, 1):
        a, ia, ma = oracle.pagerank(g, sink_mode=sink_mode)
        b, ib, mb = oracle.pagerank(g, sink_mode=sink_mode, teleport=np.full(z, 1.0 / z))
        assert ia == ib and np.abs(a - b).max() <= 1e-15


def test_oracle_teleport_linear():
    # the unnormalised fixed point is linear in the teleport vector
    z, u, v, w = random_graph(800, 3000, 9, directed=True)
    g = oracle.in_lists(z, u, v, w, True)
    t1 = _teleport(z, [1, 5, 9])
    t2 = _teleport(z, [100, 200])
    x1, *_ = solve_fixed_point(g, 0.85, 0, teleport=0.15 * t1)
    x2, *_ = solve_fixed_point(g, 0.85, 0, teleport=0.15 * t2)
    t = 0.3 * t1 + 0.7 * t2
    s, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=5000, teleport=t)
    mix = 0.3 * x1 + 0.7 * x2
    assert np.abs(s - mix / mix.sum()).sum() < 1e-11


# ---------------------------------------------------------------- eigenvector and Katz
EIGHT_EV = [0.2254, 0.1503, 0.5290, 0.4508, 0.3006, 0.5290, 0.2254, 0.1503]  # :570-577, 1e-4


def adjacency(g):
    """Dense A over existing nodes (compact order): A[u, v] = sum of the weights of u's
    in-list entries with source v, as the reference loops pull them."""
    exists = np.ones(g.z, bool) if g.exists is None else g.exists.astype(bool)
    idx = -np.ones(g.z, np.int64)
    idx[exists] = np.arange(int(exists.sum()))
    rows = np.repeat(np.arange(g.z), np.diff(g.in_ptr.astype(np.int64)))
    w = np.ones(len(rows)) if g.in_w is None else g.in_w
    n = int(exists.sum())
    A = np.zeros((n, n))
    np.add.at(A, (idx[rows], idx[g.in_src.astype(np.int64)]), w)
    return A, exists


def test_eigenvector_eight_node_known_answers():
    z, u, v, w = eight_node()
    s, it = oracle.eigenvector(oracle.in_lists(z, u, v, w, False))
    assert np.allclose(s, EIGHT_EV, atol=1e-4)
    assert it > 1


@pytest.mark.parametrize("directed", [False, True])
def test_eigenvector_is_the_leading_eigenvector(directed):
    """Dense eigendecomposition: the Perron vector of A normalised in the 2-norm.  A transposed
    pull (out-lists instead of in-lists) fails the directed case."""
    z, u, v, w = random_graph(80, 900, 61, directed=directed, weighted=True)
    g = oracle.in_lists(z, u, v, w, directed)
    A, _ = adjacency(g)
    vals, vecs = np.linalg.eig(A)
    k = int(np.argmax(vals.real))
    ref = np.abs(vecs[:, k].real)
    ref /= np.linalg.norm(ref)
    s, it = oracle.eigenvector(g, tol=1e-15, max_iterations=100000)
    assert np.abs(s - ref).max() <= 1e-6
    gt = oracle.in_lists(z, v, u, w, True) if directed else None
    if directed:
        st, _ = oracle.eigenvector(gt, tol=1e-15, max_iterations=100000)
        assert np.abs(st - ref).max() > 1e-3


def test_eigenvector_regular_graph_deleted_slot_and_sign():
    # a 7-cycle plus chords of a circulant (4-regular): uniform 1/sqrt(n); a deleted slot keeps 1
    n = 9
    u = np.concatenate([np.arange(7), np.arange(7)])
    v = np.concatenate([(np.arange(7) + 1) % 7, (np.arange(7) + 3) % 7])
    ex = np.ones(n, np.uint8)
    ex[[7, 8]] = 0
    g = oracle.in_lists(n, u, v, None, False, exists=ex)
    s, _ = oracle.eigenvector(g)
    assert np.allclose(s[:7], 1 / np.sqrt(7), atol=1e-12)
    assert s[7] == 1.0 and s[8] == 1.0
    # negative weights: the loop corrects the sign when slot 0 ends negative
    z, uu, vv, w = random_graph(40, 300, 62, directed=False, weighted=True)
    gn = oracle.in_lists(z, uu, vv, -w, False)
    sn, _ = oracle.eigenvector(gn, tol=1e-12)
    assert sn[0] >= 0


def test_katz_wiki_vote_known_answer():
    # CentralityGTest.cpp:216-225 reads the file with SNAPGraphReader{}, whose default is
    # undirected (SNAPGraphReader.hpp:42), despite the test's name
    z, u, v = wiki_vote(False)
    g = oracle.in_lists(z, u, v, None, False)
    s, it = oracle.katz(g, alpha=5e-4)
    assert oracle.ranking(s)[0] == 699
    assert it > 2


@pytest.mark.parametrize("directed", [False, True])
def test_katz_is_the_linear_solve(directed):
    """x = alpha A (1 + x) + beta has the fixed point (I - alpha A) x = alpha A 1 + beta 1; the
    result is x / ||x||_2.  A dropped (1 + .) or a missing beta fails it."""
    z, u, v, w = random_graph(500, 4000, 63, directed=directed, weighted=True)
    g = oracle.in_lists(z, u, v, w, directed)
    A, _ = adjacency(g)
    alpha = oracle.katz_default_alpha(g)
    beta = 0.3
    x = np.linalg.solve(np.eye(len(A)) - alpha * A, alpha * A.sum(1) + beta)
    ref = x / np.linalg.norm(x)
    s, it = oracle.katz(g, alpha=alpha, beta=beta, tol=1e-14)
    assert np.abs(s - ref).max() <= 1e-10


def test_katz_default_alpha_and_regular_graph():
    z, u, v, w = random_graph(300, 2000, 64, directed=True)
    g = oracle.in_lists(z, u, v, None, True)
    outdeg = np.bincount(u, minlength=z)
    assert oracle.katz_default_alpha(g) == 1.0 / (1 + outdeg.max())
    gu = oracle.in_lists(z, u, v, None, False)
    _, cnt = oracle.out_degrees(gu)
    assert oracle.katz_default_alpha(gu) == 1.0 / (1 + cnt.max())
    # 4-regular circulant: uniform scores
    n = 11
    uu = np.concatenate([np.arange(n), np.arange(n)])
    vv = np.concatenate([(np.arange(n) + 1) % n, (np.arange(n) + 2) % n])
    s, _ = oracle.katz(oracle.in_lists(n, uu, vv, None, False))
    assert np.allclose(s, 1 / np.sqrt(n), atol=1e-12)


# ---- AlgebraicPageRank and the CSR product (AlgebraicPageRank.hpp:35-119,
#      CSRGeneralMatrix.hpp:151-182, 563-578) ----

def test_for_edges_order():
    # directed: u ascending, each out-list in insertion order
    eu, ev, ew = oracle.for_edges(4, [2, 0, 2, 1, 0], [1, 3, 0, 1, 2], [1., 2., 3., 4., 5.])
    assert eu.tolist() == [0, 0, 1, 2, 2]
    assert ev.tolist() == [3, 2, 1, 1, 0]
    assert ew.tolist() == [2., 5., 4., 1., 3.]
    # undirected: each edge once, at its endpoint u >= v; self-loop once
    eu, ev, _ = oracle.for_edges(4, [0, 3, 2, 1], [2, 1, 2, 3], None, directed=False)
    assert list(zip(eu.tolist(), ev.tolist())) == [(2, 0), (2, 2), (3, 1), (3, 1)]


def test_csr_spmv_matches_scipy():
    rng = np.random.default_rng(3)
    rows, cols, nnz = 300, 200, 4000
    r = np.sort(rng.integers(0, rows, nnz))
    row_idx = np.zeros(rows + 1, np.uint64)
    row_idx[1:] = np.cumsum(np.bincount(r, minlength=rows))
    a = oracle.CSR(rows, cols, row_idx, rng.integers(0, cols, nnz), rng.normal(size=nnz))
    x = rng.normal(size=cols)
    # duplicates inside a row stay separate entries and still add up
    np.testing.assert_allclose(oracle.csr_spmv(a, x), a.to_scipy() @ x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("directed", [True, False])
def test_algebraic_matrix_is_d_At_over_outdeg(directed):
    z, u, v, w = random_graph(40, 300, 8, directed=directed, weighted=True)
    eu, ev, ew = oracle.for_edges(z, u, v, w, directed)
    mat = oracle.algebraic_matrix(z, eu, ev, ew, directed, damp=0.85)
    A = np.zeros((z, z))
    for a, b, c in zip(u, v, w):
        A[a, b] += c
        if not directed and a != b:
            A[b, a] += c
    deg = A.sum(axis=1)
    inv = np.where(deg > 0, 1.0 / np.where(deg > 0, deg, 1.0), 0.0)
    M = 0.85 * (A * inv[:, None]).T
    np.testing.assert_allclose(mat.to_scipy().toarray(), M, rtol=1e-14, atol=1e-15)
    assert mat.nnz == (len(u) if directed else len(u) + int((u != v).sum()))


def test_algebraic_pagerank_known_answers():
    # AlgebraicPageRankGTest.cpp:55-94 (Matlab values, +-1e-4)
    z, u, v, w = eight_node()
    eu, ev, ew = oracle.for_edges(z, u, v, w, directed=False)
    s, _ = oracle.algebraic_pagerank(oracle.algebraic_matrix(z, eu, ev, ew, False))
    np.testing.assert_allclose(np.abs(s), EIGHT_PR, atol=1e-4)
    # AlgebraicPageRankGTest.cpp:21-52: wiki-Vote as SNAPGraphReader reads it (undirected),
    # top-1 node 699 with 0.00432 +- 1e-3
    z, u, v = wiki_vote(False)
    eu, ev, _ = oracle.for_edges(z, u, v, None, directed=False)
    s, _ = oracle.algebraic_pagerank(oracle.algebraic_matrix(z, eu, ev, None, False))
    top = int(np.lexsort((np.arange(z), -s))[0])
    assert top == WIKI_TOP[False][0]
    assert abs(s[top] - WIKI_TOP[False][1]) <= 1e-3


@pytest.mark.parametrize("directed", [True, False])
def test_algebraic_pagerank_fixed_point_and_pagerank(directed):
    """The fixed point is the solution of (I - M) x = (1-d)/n 1 normalised; with every slot
    present it is also PageRank's result (the same loop, M folded), same iteration count."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    z, u, v, w = random_graph(500, 4000, 21, directed=directed, weighted=True)
    eu, ev, ew = oracle.for_edges(z, u, v, w, directed)
    mat = oracle.algebraic_matrix(z, eu, ev, ew, directed)
    s, it = oracle.algebraic_pagerank(mat, tol=1e-13)
    x = spl.spsolve(sp.identity(z, format="csc") - mat.to_scipy().tocsc(),
                    np.full(z, 0.15 / z))
    np.testing.assert_allclose(s, x / x.sum(), rtol=0, atol=1e-12)
    s8, it8 = oracle.algebraic_pagerank(mat)
    g = oracle.in_lists(z, u, v, w, directed)
    ref, it_ref, _ = oracle.pagerank(g)
    assert it8 == it_ref
    assert np.abs(s8 - ref).sum() <= 1e-12
    # the cap stops the loop
    assert oracle.algebraic_pagerank(mat, max_iterations=3)[1] == 3


# ---- ApproximatePageRank and PageRankNibble's sweep (ApproximatePageRank.cpp:18-61,
#      PageRankNibble.cpp:28-85) ----

def lazy_ppr(z, ptr, dst, w, s, alpha):
    """pr_alpha(s): the solution of pr = alpha s + (1 - alpha) W^T pr with the lazy walk
    W = (I + D^-1 A') / 2 (sparse direct solve), where A' is the adjacency without its
    diagonal and D the weighted degree with self-loops counted twice: a push sends nothing
    through a self-loop (the reference's r[u] assignment drops that share), and on graphs
    without self-loops W is the usual lazy walk (I + D^-1 A) / 2."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    rows = np.repeat(np.arange(z), np.diff(np.asarray(ptr, np.int64)))
    cols = np.asarray(dst, np.int64)
    wt = np.ones(len(dst)) if w is None else np.asarray(w, np.float64)
    off = rows != cols
    A = sp.csr_matrix((wt[off], (rows[off], cols[off])), shape=(z, z))
    d = np.bincount(rows, weights=np.where(off, wt, 2.0 * wt), minlength=z)
    # a node without out-edges sends nothing: its row is the lazy half alone
    inv = np.divide(1.0, d, out=np.zeros(z), where=d > 0)
    W = 0.5 * (sp.identity(z) + sp.diags(inv) @ A)
    M = sp.identity(z, format="csc") - (1 - alpha) * W.T.tocsc()
    return alpha * spl.spsolve(M, s)


@pytest.mark.parametrize("self_loops", [False, True])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("seeds", [[5], [3, 17, 40]])
def test_apr_acl_identity_and_invariants(weighted, seeds, self_loops):
    """p + pr_alpha(r) = pr_alpha(seeds) exactly (any push order keeps it), every residual
    below eps vol, mass conserved, p >= 0, touched = support of p and r.  With self-loops the
    reference drops the share a push sends through one: the identity holds for the walk
    without them (and fails for the walk that keeps them), and mass leaks."""
    z, u, v, w = random_graph(400, 2400, 9, directed=False, weighted=weighted,
                              self_loops=self_loops)
    if self_loops:                           # loops at the seeds and around them
        extra = np.array([5, 3, 17, 40, 6, 7, 8, 9])
        u, v = np.concatenate([u, extra]), np.concatenate([v, extra])
        if w is not None:
            w = np.concatenate([w, np.full(len(extra), 0.75)])
    ptr, dst, ww = oracle.out_lists(z, u, v, w, directed=False)
    vol = oracle.weighted_degree2(z, ptr, dst, ww)
    assert (vol > 0).all()
    alpha, eps = 0.1, 1e-7
    p, r, touched = oracle.approximate_pagerank(z, ptr, dst, ww, seeds, alpha, eps)
    s = np.zeros(z)
    s[np.unique(seeds)] = 1.0 / len(np.unique(seeds))
    np.testing.assert_allclose(p + lazy_ppr(z, ptr, dst, ww, r, alpha),
                               lazy_ppr(z, ptr, dst, ww, s, alpha), rtol=0, atol=1e-12)
    assert (r < eps * vol).all()
    if self_loops:
        assert p.sum() + r.sum() < 1.0 - 1e-4
        # the walk that keeps the self-loop share (D^-1 A with A's diagonal) is not the one
        kept = sp_lazy_keep(z, ptr, dst, ww, alpha)
        assert np.abs(p + kept(r) - kept(s)).max() > 1e-6
    else:
        assert abs(p.sum() + r.sum() - 1.0) <= 1e-12
    assert (p >= 0).all()
    assert np.array_equal(touched.astype(bool), (p > 0) | (r > 0))


@pytest.mark.parametrize("weighted", [False, True])
def test_apr_directed_with_sinks(weighted):
    """Directed graphs (forNeighborsOf = out-edges): a node without out-edges is never queued by
    a neighbour's add, so its residual stays (the reference's r[v] < vol eps is false at vol
    0), and a seed without out-edges is pushed until its residual is gone (r / 0 >= eps),
    losing the half of every push it cannot send.  The identity holds with the lazy walk whose
    rows at such nodes are the lazy half alone; every other push conserves mass; residuals at
    nodes with out-edges end below eps vol."""
    z, u, v, w = random_graph(500, 1500, 17, directed=True, weighted=weighted, self_loops=False)
    ptr, dst, ww = oracle.out_lists(z, u, v, w, directed=True)
    vol = oracle.weighted_degree2(z, ptr, dst, ww)
    sinks = np.nonzero(vol == 0)[0]
    assert len(sinks) > 10
    seeds = [int(np.nonzero(vol > 0)[0][0]), int(sinks[0])]
    alpha, eps = 0.15, 1e-8
    p, r, touched = oracle.approximate_pagerank(z, ptr, dst, ww, seeds, alpha, eps)
    s = np.zeros(z)
    s[np.unique(seeds)] = 0.5
    np.testing.assert_allclose(p + lazy_ppr(z, ptr, dst, ww, r, alpha),
                               lazy_ppr(z, ptr, dst, ww, s, alpha), rtol=0, atol=1e-12)
    # every push of the sink seed adds alpha res to p and has nowhere to send (1 - alpha) res / 2:
    # (1 - alpha) / (2 alpha) p[seed] is lost in all; every other push conserves mass
    lost = (1 - alpha) / (2 * alpha) * p[seeds[1]]
    assert lost > 0.3 and abs(p.sum() + r.sum() - (1.0 - lost)) <= 1e-12
    assert (r[vol > 0] < eps * vol[vol > 0]).all()
    assert r[seeds[1]] == 0.0
    assert r[sinks].sum() > 0                 # mass that reached sinks stays there


def sp_lazy_keep(z, ptr, dst, w, alpha):
    """pr_alpha for the lazy walk that keeps a self-loop's share: (I + D^-1 A) / 2, A with its
    diagonal counted twice like D (a planted alternative the reference does not follow)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    rows = np.repeat(np.arange(z), np.diff(np.asarray(ptr, np.int64)))
    cols = np.asarray(dst, np.int64)
    wt = np.ones(len(dst)) if w is None else np.asarray(w, np.float64)
    wt = np.where(rows == cols, 2.0 * wt, wt)
    A = sp.csr_matrix((wt, (rows, cols)), shape=(z, z))
    d = np.asarray(A.sum(axis=1)).ravel()
    M = sp.identity(z, format="csc") - (1 - alpha) * (0.5 * (sp.identity(z) + sp.diags(1.0 / d) @ A)).T.tocsc()
    return lambda x: alpha * spl.spsolve(M, x)


def test_sweep_matches_brute_force():
    z, u, v, w = random_graph(300, 1500, 10, directed=False, weighted=True, self_loops=True)
    ptr, dst, ww = oracle.out_lists(z, u, v, w, directed=False)
    rng = np.random.default_rng(2)
    ids = rng.choice(z, 60, replace=False).astype(np.uint64)
    score = rng.random(60)
    score[5] = score[6]                      # a tie, ordered by id
    total = 2 * float(np.asarray(w).sum())   # self-loops once, like totalEdgeWeight
    order, best, cond = oracle.best_sweep_set(z, ptr, dst, ww, z, total, ids, score)
    want = ids[np.lexsort((ids, -score))].astype(np.int64)
    assert np.array_equal(order, want)
    pi = np.asarray(ptr, np.int64)
    conds = []
    for L in range(1, len(order) + 1):
        S = set(order[:L].tolist())
        vol = cut = 0.0
        for x in S:
            for e in range(pi[x], pi[x + 1]):
                n = int(dst[e])
                vol += ww[e]
                if n not in S or n == x:
                    cut += ww[e]
        conds.append(cut / min(vol, total - vol))
    k = int(np.argmin(conds))
    assert best == k + 1
    assert abs(cond - conds[k]) <= 1e-9 * abs(conds[k])


def test_apr_reference_inputs():
    """test_scd.py:12-25 and SelectiveCDGTest.cpp:33-39: non-empty results with non-negative
    scores of existing nodes, the same for a seed and the one-seed list."""
    from nkbg_inputs import metis
    for name, seed, alpha, eps in (("PGPgiantcompo.graph", 0, 0.1, 1e-12),
                                   ("hep-th.graph", 50, 0.1, 1e-5)):
        n, u, v = metis(name)
        ptr, dst, ww = oracle.out_lists(n, u, v, None, directed=False)
        p, r, touched = oracle.approximate_pagerank(n, ptr, dst, ww, [seed], alpha, eps)
        assert touched.sum() >= 1 and (p >= 0).all() and p[seed] > 0
        assert abs(p.sum() + r.sum() - 1.0) <= 1e-9
