"""CSR products and AlgebraicPageRank on the GPU against the oracle
(AlgebraicPageRank.hpp:35-119, CSRGeneralMatrix.hpp:151-182, 563-578).

  nkpr_csr_spmv: row length mixes that select every lane width (4..64 lanes per row) and the
      long-row path (rows longer than 32 x the lanes per row, cut into pieces of 4096
      entries), with empty rows and duplicate entries, against oracle.csr_spmv (sums in
      another order: 1e-13 relative to the sum of |terms|);
  nkpr_csr_pagerank_step: y and ||y - x||^2 against numpy;
  nkpr_algebraic_pagerank and the handle route (nkpr_algebraic_create/run, which the C++/Cython
      drop-in takes): the oracle's iteration count and L1 <= 1e-12 on the reference's
      known-answer graphs, graphs with deleted slots and skewed graphs with long rows; results
      bit-identical between runs.
"""
import numpy as np
import pytest

import oracle
from graphs import EIGHT_PR, WIKI_TOP, eight_node, random_graph, skewed_graph, wiki_vote
from networkit_amd import capi

pytestmark = pytest.mark.gpu


def random_csr(rows, cols, lengths, seed):
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, np.int64)
    row_idx = np.zeros(rows + 1, np.uint64)
    row_idx[1:] = np.cumsum(lengths)
    nnz = int(row_idx[-1])
    col = rng.integers(0, cols, nnz).astype(np.uint64)
    # duplicates inside rows: repeat some columns
    if nnz > 10:
        col[1::7] = col[0:-1:7][: len(col[1::7])]
    val = rng.normal(size=nnz)
    return oracle.CSR(rows, cols, row_idx, col, val)


def check_spmv(a, seed=0):
    import torch
    rng = np.random.default_rng(seed)
    x = rng.normal(size=a.cols)
    ref = oracle.csr_spmv(a, x)
    d = capi.DeviceCSR(a.rows, a.cols, a.row_idx, a.col_idx, a.values)
    y = d.spmv(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    got = y.cpu().numpy()
    scale = oracle.csr_spmv(oracle.CSR(a.rows, a.cols, a.row_idx, a.col_idx, np.abs(a.values)),
                            np.abs(x))
    assert np.all(np.abs(got - ref) <= 1e-13 * scale + 1e-300), np.max(np.abs(got - ref))
    return d, x, got


@pytest.mark.parametrize("avg", [2, 9, 20, 40, 120])
def test_csr_spmv_lane_widths(avg):
    rng = np.random.default_rng(avg)
    rows = 20000
    lengths = rng.poisson(avg, rows)
    lengths[::97] = 0                       # empty rows
    check_spmv(random_csr(rows, 15000, lengths, avg), seed=avg)


def test_csr_spmv_long_rows():
    rng = np.random.default_rng(5)
    rows = 5000
    lengths = rng.poisson(6, rows)
    long_rows = rng.choice(rows, 12, replace=False)
    lengths[long_rows] = rng.integers(2049, 60000, 12)
    lengths[long_rows[0]] = 2048            # exactly the short-row limit
    lengths[long_rows[1]] = 2049            # the first long length
    lengths[rows - 1] = 70000               # a long last row
    check_spmv(random_csr(rows, 9000, lengths, 6), seed=7)


def test_csr_spmv_rectangular_and_tiny():
    check_spmv(random_csr(1, 5, [5], 1))
    check_spmv(random_csr(3, 1, [0, 0, 4], 2))
    check_spmv(random_csr(700, 50, np.full(700, 3), 3))


def test_csr_pagerank_step():
    import torch
    rng = np.random.default_rng(11)
    n = 30000
    lengths = rng.poisson(12, n)
    lengths[:5] = 5000
    a = random_csr(n, n, lengths, 12)
    x = rng.random(n)
    y_ref = oracle.csr_spmv(a, x) + 0.25
    d = capi.DeviceCSR(n, n, a.row_idx, a.col_idx, a.values)
    tx = torch.from_numpy(x).cuda()
    y, sq = d.pagerank_step(tx, 0.25)
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.cpu().numpy(), y_ref, rtol=1e-12, atol=1e-12)
    ref_sq = float(((y_ref - x) ** 2).sum())
    assert abs(float(sq.item()) - ref_sq) <= 1e-12 * ref_sq
    # the same call again gives the same bits (partials added in a fixed order)
    y2, sq2 = d.pagerank_step(tx, 0.25)
    torch.cuda.synchronize()
    assert torch.equal(y, y2) and float(sq.item()) == float(sq2.item())


def graph_cases():
    yield "eight", eight_node(), False
    z, u, v = wiki_vote(False)
    yield "wiki_undirected", (z, u, v, None), False
    z, u, v = wiki_vote(True)
    yield "wiki_directed", (z, u, v, None), True
    yield "random_weighted", random_graph(3000, 30000, 5, directed=True, weighted=True), True
    yield "skewed_long_rows", skewed_graph(20000, 400000, 9), True


@pytest.mark.parametrize("case", [c[0] for c in graph_cases()])
def test_algebraic_pagerank_capi(case):
    name, (z, u, v, w), directed = next(c for c in graph_cases() if c[0] == case)
    eu, ev, ew = oracle.for_edges(z, u, v, w, directed)
    mat = oracle.algebraic_matrix(z, eu, ev, ew, directed)
    ref, it = oracle.algebraic_pagerank(mat)
    d = capi.DeviceCSR(z, z, mat.row_idx, mat.col_idx, mat.values)
    got, it_dev = d.algebraic_pagerank()
    e = np.abs(got - ref).sum()
    print("%s: %d iterations (oracle %d), L1 %.3e" % (name, it_dev, it, e), flush=True)
    assert it_dev == it
    assert e <= 1e-12
    again, _ = d.algebraic_pagerank()
    assert np.array_equal(got, again)
    if name == "eight":
        np.testing.assert_allclose(got, EIGHT_PR, atol=1e-4)
    if name == "wiki_undirected":
        top = int(np.lexsort((np.arange(z), -got))[0])
        assert top == WIKI_TOP[False][0] and abs(got[top] - WIKI_TOP[False][1]) <= 1e-3
    if name == "skewed_long_rows":
        assert np.diff(mat.row_idx.astype(np.int64)).max() > 2048


@pytest.mark.parametrize("case", [c[0] for c in graph_cases()])
@pytest.mark.parametrize("binned", ["0", "1"])
def test_algebraic_handle(monkeypatch, case, binned):
    """nkpr_algebraic_create/run: the PageRank handle built from M (in-lists = M's rows,
    coefficients = M's values) runs the reference's loop with the oracle's iteration count and
    L1 <= 1e-12, through the pull sweep and through the binned FP64 copy (forced here, the
    default from 2^26 entries); a second run on the same handle repeats the bits."""
    monkeypatch.setenv("NKPR_BINNED64", binned)
    name, (z, u, v, w), directed = next(c for c in graph_cases() if c[0] == case)
    eu, ev, ew = oracle.for_edges(z, u, v, w, directed)
    mat = oracle.algebraic_matrix(z, eu, ev, ew, directed)
    ref, it = oracle.algebraic_pagerank(mat)
    d = capi.DeviceCSR(z, z, mat.row_idx, mat.col_idx, mat.values)
    h = capi.DeviceAlgebraic(d)
    got, it_dev = h.run()
    again, it2 = h.run()
    h.close()
    e = np.abs(got - ref).sum()
    print("%s (binned %s): %d iterations (oracle %d), L1 %.3e" % (name, binned, it_dev, it, e),
          flush=True)
    assert it_dev == it == it2
    assert e <= 1e-12
    assert np.array_equal(got, again)
    if name == "eight":
        np.testing.assert_allclose(got, EIGHT_PR, atol=1e-4)


def test_algebraic_handle_rejects_malformed_matrix():
    a = random_csr(100, 100, np.full(100, 4), 22)
    col = a.col_idx.copy()
    col[5] = 1000
    d = capi.DeviceCSR(100, 100, a.row_idx, col, a.values)
    with pytest.raises(capi.NkprError, match="malformed"):
        capi.DeviceAlgebraic(d)
    rect = capi.DeviceCSR(100, 50, a.row_idx, a.col_idx % 50, a.values)
    with pytest.raises(capi.NkprError, match="square"):
        capi.DeviceAlgebraic(rect)


def test_algebraic_pagerank_drop_in_with_deleted_nodes():
    """The C++ class builds M from the Graph like the reference (forEdges order, deleted slots
    are empty rows that still get the teleport term) and runs on the device."""
    import networkit_amd as nk
    z, u, v, w = random_graph(500, 5000, 13, directed=True, weighted=True)
    G = nk.Graph(z, True, True)
    G.addEdges(u, v, w)
    for x in (3, 77, 499):
        G.removeNode(x)
    apr = nk.algebraic.AlgebraicPageRank(G, 0.85, 1e-10)
    apr.run()
    got = np.array(apr.scores())
    keep = np.isin(u, [3, 77, 499], invert=True) & np.isin(v, [3, 77, 499], invert=True)
    # removeNode drops incident edges; the oracle sees the remaining edges in the order the
    # Graph keeps them (out-lists after removal keep the insertion order of what remains for
    # edges that were never swapped; compare through the fixed point, order-free)
    eu, ev, ew = oracle.for_edges(z, u[keep], v[keep], w[keep], True)
    mat = oracle.algebraic_matrix(z, eu, ev, ew, True)
    ref, it = oracle.algebraic_pagerank(mat, tol=1e-10)
    assert np.abs(got - ref).sum() <= 1e-12
    assert abs(apr.numberOfIterations() - it) <= 1
    assert apr.maximum() == 1.0
    rk = apr.ranking()
    assert len(rk) == z
    assert all(rk[i][1] >= rk[i + 1][1] for i in range(len(rk) - 1))


def test_algebraic_pagerank_rejects_malformed_matrix():
    """A column index >= cols is found by the structure check before any product runs."""
    a = random_csr(100, 100, np.full(100, 4), 21)
    col = a.col_idx.copy()
    col[17] = 100
    d = capi.DeviceCSR(100, 100, a.row_idx, col, a.values)
    with pytest.raises(capi.NkprError, match="malformed"):
        d.algebraic_pagerank()


def test_csr_spmv_many_long_rows_and_pieces():
    """Power-law rows: thousands of rows above the short-row limit spread over many scan
    blocks, rows of several pieces (4096 entries each), against the oracle product; the step on
    the same matrix is bit-identical on a second call."""
    import torch
    rng = np.random.default_rng(31)
    rows = 400_000
    lengths = np.minimum((rng.pareto(1.2, rows) * 3).astype(np.int64), 40_000)
    a = random_csr(rows, rows, lengths, 32)
    d, x, got = check_spmv(a, seed=33)
    lens = np.diff(a.row_idx.astype(np.int64))
    assert (lens > 32 * 16).sum() > 500 and (lens > 4096).sum() > 20
    tx = torch.from_numpy(x).cuda()
    y1, s1 = d.pagerank_step(tx, 0.1)
    y2, s2 = d.pagerank_step(tx, 0.1)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2) and float(s1.item()) == float(s2.item())
