"""GPU parity of PageRank on a batch of graphs (nkpr_graph_create_batch / nkpr_run_batch).

Every graph of the batch must come out exactly as if it had been run alone:
  FP64, reference stopping rule: the oracle's iteration count for that graph and an L1
      difference <= 1e-12 per graph; maximum() equal to 1e-12 relative.
  MIXED: every graph certified, L1 error <= tol against the converged oracle.
  FP32 with the certified stop at 1e-5: L1 error <= 1e-5 + 2e-6 per graph.
"""
import numpy as np
import pytest

import oracle
from graphs import random_graph, skewed_graph
from networkit_amd import capi

pytestmark = pytest.mark.gpu


def make_batch(specs, weighted, seed):
    """specs: list of (n, m, directed, dead slots).  Returns the per-graph oracle in-lists and the
    union CSC arrays."""
    rng = np.random.default_rng(seed)
    graphs, ptrs, srcs, ws, ex, offs = [], [np.zeros(1, np.uint64)], [], [], [], [0]
    base_e = 0
    for i, (n, m, directed, dead) in enumerate(specs):
        if i % 3 == 2:
            z, u, v, w = skewed_graph(n, m, int(rng.integers(1 << 30)), directed=directed,
                                      weighted=weighted)
        else:
            z, u, v, w = random_graph(n, m, int(rng.integers(1 << 30)), directed=directed,
                                      weighted=weighted)
        exists = None
        if dead:
            exists = np.ones(z, np.uint8)
            exists[list(dead)] = 0
            keep = ~np.isin(u, dead) & ~np.isin(v, dead)
            u, v = u[keep], v[keep]
            w = None if w is None else w[keep]
        g = oracle.in_lists(z, u, v, w, directed, exists=exists)
        graphs.append(g)
        ptrs.append(g.in_ptr[1:] + np.uint64(base_e))
        srcs.append(g.in_src.astype(np.int64) + offs[-1])
        if weighted:
            ws.append(g.in_w)
        ex.append(np.ones(z, np.uint8) if exists is None else exists)
        base_e += g.m
        offs.append(offs[-1] + z)
    colptr = np.concatenate(ptrs).astype(np.uint64)
    src = np.concatenate(srcs).astype(np.int32)
    w = np.concatenate(ws) if weighted else None
    exists = np.concatenate(ex)
    return graphs, offs, colptr, src, w, exists


def run_batch(offs, colptr, src, w, exists, **kw):
    dg = capi.DeviceGraph(offs[-1], colptr, src, None if w is None else w.astype(np.float32),
                          exists=exists, weight64=w, graph_offsets=offs)
    x = np.full(offs[-1], -1.0)
    for q in range(len(offs) - 1):
        n = int(exists[offs[q]:offs[q + 1]].sum())
        x[offs[q]:offs[q + 1]] = 1.0 / max(n, 1)
    st, its, mx = dg.run_batch(x, capi.default_params(**kw))
    return x, st, its, mx


SPECS = [(800, 3000, True, ()), (2500, 9000, True, (3, 70)), (60, 150, True, ()),
         (4000, 20000, False, ()), (1, 0, True, ()), (1500, 1500, True, ()),
         (3000, 12000, False, (0,)), (700, 5000, True, ()), (2, 1, True, ())]


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("norm,sinks,normalized", [(1, 0, False), (0, 1, False), (1, 0, True)])
def test_batch_graphs_fp64_parity(weighted, norm, sinks, normalized):
    graphs, offs, colptr, src, w, exists = make_batch(SPECS, weighted, 11)
    x, st, its, mx = run_batch(offs, colptr, src, w, exists, norm=norm, sink_handling=sinks,
                               normalized=int(normalized))
    for q, g in enumerate(graphs):
        ref, it, rmx = oracle.pagerank(g, norm=norm, sink_mode=sinks, normalized=normalized)
        got = x[offs[q]:offs[q + 1]]
        assert its[q] == it, (q, its[q], it)
        scale = 1.0 if not normalized else max(1.0, np.abs(ref).max())
        ex = np.ones(g.z, bool) if g.exists is None else g.exists.astype(bool)
        assert np.abs(got[ex] - ref[ex]).sum() <= 1e-12 * scale, q
        assert mx[q] == pytest.approx(rmx, rel=1e-12)
    assert st["iterations"] == max(its)


def test_batch_graphs_sliced_union():
    # enough nodes for the source-sliced sweep of the union (2^20 and up)
    rng = np.random.default_rng(12)
    specs = [(int(n), int(n) * 6, True, ()) for n in rng.integers(5000, 40000, 48)]
    graphs, offs, colptr, src, w, exists = make_batch(specs, True, 13)
    assert offs[-1] >= 1 << 20
    x, st, its, mx = run_batch(offs, colptr, src, w, exists)
    for q, g in enumerate(graphs):
        ref, it, _ = oracle.pagerank(g)
        assert its[q] == it, q
        assert np.abs(x[offs[q]:offs[q + 1]] - ref).sum() <= 1e-12, q


@pytest.mark.parametrize("sinks", [0, 1])
def test_batch_graphs_mixed_certified(sinks):
    graphs, offs, colptr, src, w, exists = make_batch(SPECS, True, 14)
    x, st, its, mx = run_batch(offs, colptr, src, w, exists, precision=capi.NKPR_MIXED,
                               tol=1e-9, sink_handling=sinks)
    assert st["l1_error_bound"] <= 1e-9
    for q, g in enumerate(graphs):
        ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000,
                                    sink_mode=sinks)
        ex = np.ones(g.z, bool) if g.exists is None else g.exists.astype(bool)
        assert np.abs(x[offs[q]:offs[q + 1]][ex] - ref[ex]).sum() <= 1e-9, q


def test_batch_graphs_fp32_certified():
    graphs, offs, colptr, src, w, exists = make_batch(SPECS, True, 15)
    x, st, its, mx = run_batch(offs, colptr, src, w, exists, precision=capi.NKPR_FP32,
                               stop=capi.NKPR_STOP_L1_ERROR, tol=1e-5)
    for q, g in enumerate(graphs):
        ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000)
        ex = np.ones(g.z, bool) if g.exists is None else g.exists.astype(bool)
        assert np.abs(x[offs[q]:offs[q + 1]][ex] - ref[ex]).sum() <= 1e-5 + 2e-6, q


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("sinks", [0, 1])
def test_batch_graphs_binned_copies(monkeypatch, weighted, sinks):
    """Both binned copies forced on a batch handle (the default on unions of 2^26+ edges): the
    FP64 run keeps every graph's oracle iteration count and L1 <= 1e-12 with deleted slots and
    both sink rules, and the mixed run (binned GMRES operator, binned residual sweeps)
    certifies every graph."""
    monkeypatch.setenv("NKPR_BINNED", "1")
    monkeypatch.setenv("NKPR_BINNED64", "1")
    graphs, offs, colptr, src, w, exists = make_batch(SPECS, weighted, 17)
    x, st, its, mx = run_batch(offs, colptr, src, w, exists, norm=oracle.L1, sink_handling=sinks)
    for q, g in enumerate(graphs):
        ref, it, rmx = oracle.pagerank(g, norm=oracle.L1, sink_mode=sinks)
        ex = np.ones(g.z, bool) if g.exists is None else g.exists.astype(bool)
        assert its[q] == it, (q, its[q], it)
        assert np.abs(x[offs[q]:offs[q + 1]][ex] - ref[ex]).sum() <= 1e-12, q
        assert mx[q] == pytest.approx(rmx, rel=1e-12)
    x, st, its, mx = run_batch(offs, colptr, src, w, exists, precision=capi.NKPR_MIXED,
                               tol=1e-9, sink_handling=sinks)
    assert st["l1_error_bound"] <= 1e-9
    for q, g in enumerate(graphs):
        ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000,
                                    sink_mode=sinks)
        ex = np.ones(g.z, bool) if g.exists is None else g.exists.astype(bool)
        assert np.abs(x[offs[q]:offs[q + 1]][ex] - ref[ex]).sum() <= 1e-9, q


@pytest.mark.parametrize("item", ["0", "3000"])
def test_batch_graphs_binned_union_spans_chunks(monkeypatch, item):
    """A union of 48 graphs over many source chunks and destination bins (item=3000 splits bins
    over work items) through both forced binned copies: FP64 per-graph parity with the oracle
    and a certified mixed run."""
    monkeypatch.setenv("NKPR_BINNED", "1")
    monkeypatch.setenv("NKPR_BINNED64", "1")
    monkeypatch.setenv("NKPR_BINNED_ITEM", item)
    rng = np.random.default_rng(12)
    specs = [(int(n), int(n) * 6, True, ()) for n in rng.integers(5000, 40000, 48)]
    graphs, offs, colptr, src, w, exists = make_batch(specs, True, 13)
    x, st, its, mx = run_batch(offs, colptr, src, w, exists)
    for q, g in enumerate(graphs):
        ref, it, _ = oracle.pagerank(g)
        assert its[q] == it, q
        assert np.abs(x[offs[q]:offs[q + 1]] - ref).sum() <= 1e-12, q
    x, st, its, mx = run_batch(offs, colptr, src, w, exists, precision=capi.NKPR_MIXED,
                               tol=1e-9)
    assert st["l1_error_bound"] <= 1e-9
    for q, g in enumerate(graphs):
        ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000)
        assert np.abs(x[offs[q]:offs[q + 1]] - ref).sum() <= 1e-9, q


@pytest.mark.parametrize("sinks", [0, 1])
def test_batch_graphs_closed_form_start(monkeypatch, sinks):
    """A mixed batch whose start is constant on every graph (1/n_q) takes its first residual in
    closed form: every graph certifies and matches the converged oracle, with one FP64 sweep
    fewer than with NKPR_CLOSED_START=0; a start that varies inside a graph takes the sweep."""
    graphs, offs, colptr, src, w, exists = make_batch(SPECS, True, 18)
    refs = [oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000, sink_mode=sinks)[0]
            for g in graphs]
    runs = {}
    for flag, start in (("1", "const"), ("0", "const"), ("1", "varied")):
        monkeypatch.setenv("NKPR_CLOSED_START", flag)
        dg = capi.DeviceGraph(offs[-1], colptr, src, w.astype(np.float32), exists=exists,
                              weight64=w, graph_offsets=offs)
        x = np.full(offs[-1], -1.0)
        rng = np.random.default_rng(19)
        for q in range(len(offs) - 1):
            nq = int(exists[offs[q]:offs[q + 1]].sum())
            if start == "const":
                x[offs[q]:offs[q + 1]] = 1.0 / max(nq, 1)
            else:
                v = rng.uniform(0.5, 1.5, offs[q + 1] - offs[q])
                x[offs[q]:offs[q + 1]] = v / max(v[exists[offs[q]:offs[q + 1]] == 1].sum(), 1e-300)
        st, its, mx = dg.run_batch(x, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-9,
                                                          sink_handling=sinks))
        assert st["l1_error_bound"] <= 1e-9, (flag, start)
        for q, g in enumerate(graphs):
            ex = np.ones(g.z, bool) if g.exists is None else g.exists.astype(bool)
            assert np.abs(x[offs[q]:offs[q + 1]][ex] - refs[q][ex]).sum() <= 1e-9, (flag, start, q)
        runs[(flag, start)] = st
    closed, swept, varied = runs[("1", "const")], runs[("0", "const")], runs[("1", "varied")]
    if closed["refinements"] == swept["refinements"]:
        assert closed["sweeps_fp64"] == swept["sweeps_fp64"] - 1
    assert varied["sweeps_fp64"] == varied["refinements"] + 1


def test_batch_graphs_invalid():
    graphs, offs, colptr, src, w, exists = make_batch(SPECS[:3], False, 16)
    bad = src.copy()
    bad[0] = offs[1] + 1  # an edge of graph 0 from a node of graph 1
    with pytest.raises(capi.NkprError) as e:
        capi.DeviceGraph(offs[-1], colptr, bad, exists=exists, graph_offsets=offs)
    assert e.value.code == capi.NKPR_ERR_INVALID
    with pytest.raises(capi.NkprError):
        capi.DeviceGraph(offs[-1], colptr, src, exists=exists, graph_offsets=offs[:-1])
    dg = capi.DeviceGraph(offs[-1], colptr, src, exists=exists, graph_offsets=offs)
    with pytest.raises(capi.NkprError):
        dg.run(np.full(offs[-1], 0.1))
    assert dg.info()["n"] == int(exists.sum())
