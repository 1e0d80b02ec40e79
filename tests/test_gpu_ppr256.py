"""Batched personalised PageRank at the benchmark's column count: 256 teleport vectors, each
uniform over a seed set of U{1..1000} random nodes, on a Barabasi-Albert graph (8 attachments
per node, undirected, unweighted, as in the benchmark) and on a weighted directed graph.

Every tile width the autotuner can pick is forced in turn (NKPR_PPR_LPR = 4, 16, 64 lanes per
row: 16, 64, 256 FP32 columns or 8, 32, 128 FP64 columns per tile), in both precisions, and once
with a split threshold (NKPR_PPR_HEAVY = 256) below the graph's largest in-degree, so that rows
are cut into pieces under the 64-lane tiles.  Each column is checked against the oracle run
with that column's teleport vector:
  FP64, reference stopping rule: the oracle's iteration count and L1 <= 1e-12 per column;
  MIXED: every column certified (bound <= 1e-8) and L1 <= 1e-8 against the oracle run on to an
      L1 step of 1e-15.
"""
import numpy as np
import pytest

import oracle
from graphs import skewed_graph
from networkit_amd import capi
from networkit_amd import generators as gen

pytestmark = pytest.mark.gpu

K = 256


def teleport(z, seeds):
    t = np.zeros(z)
    np.add.at(t, np.asarray(seeds, np.int64), 1.0)
    return t / t.sum()


def make_case(kind):
    rng = np.random.default_rng(42 if kind == "ba" else 43)
    if kind == "ba":
        z = 30000
        u, v = gen.barabasi_albert(z, 8, seed=4)
        g = oracle.in_lists(z, u, v, None, False)
    else:
        z, u, v, w = skewed_graph(20000, 200000, 44)
        g = oracle.in_lists(z, u, v, w, True)
    seeds = [rng.choice(z, int(rng.integers(1, 1001)), replace=True) for _ in range(K)]
    step, tight = [], []
    for j, s in enumerate(seeds):
        if j % 64 == 0:
            print("%s: oracle column %d of %d" % (kind, j, K), flush=True)
        t = teleport(z, s)
        r, it, _ = oracle.pagerank(g, teleport=t)
        step.append((r, it))
        r2, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000,
                                   teleport=t)
        tight.append(r2)
    return g, seeds, step, tight


_CASES = {}


def case(kind):
    if kind not in _CASES:
        _CASES[kind] = make_case(kind)
    return _CASES[kind]


def device_graph(g):
    w32 = None if g.in_w is None else g.in_w.astype(np.float32)
    return capi.DeviceGraph(g.z, g.in_ptr, g.in_src.astype(np.int32), w32, weight64=g.in_w)


@pytest.mark.parametrize("kind", ["ba", "weighted"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("lpr,heavy", [(4, None), (16, None), (64, None), (64, 256)])
def test_ppr_256_columns_forced_tiles(monkeypatch, kind, precision, lpr, heavy):
    g, seeds, step, tight = case(kind)
    if heavy is not None:
        indeg = np.diff(g.in_ptr.astype(np.int64))
        assert indeg.max() > heavy  # some rows are cut into pieces
    monkeypatch.setenv("NKPR_PPR_LPR", str(lpr))
    if heavy is not None:
        monkeypatch.setenv("NKPR_PPR_HEAVY", str(heavy))
    else:
        monkeypatch.delenv("NKPR_PPR_HEAVY", raising=False)
    dg = device_graph(g)
    if precision == "fp64":
        p = capi.default_params()
    else:
        p = capi.default_params(precision=capi.NKPR_MIXED, tol=1e-8)
    got, st, its, bound = dg.ppr_batch(seeds, params=p)
    # the forced tile ran: FP64 tiles are 2 columns per lane, FP32 tiles (mixed) 4
    assert st["tile_columns"] == lpr * (2 if precision == "fp64" else 4)
    if heavy is not None:
        assert st["heavy_degree"] == heavy
    worst = 0.0
    for j in range(K):
        if precision == "fp64":
            ref, it = step[j]
            assert its[j] == it, (j, its[j], it)
            e = np.abs(got[j] - ref).sum()
            assert e <= 1e-12, (j, e)
        else:
            assert bound[j] <= 1e-8, (j, bound[j])
            e = np.abs(got[j] - tight[j]).sum()
            assert e <= 1e-8, (j, e)
        worst = max(worst, e)
    print("%s %s LPR %d heavy %s: worst column L1 %.3e, %d sweeps" % (
        kind, precision, lpr, heavy, worst, st["sweeps"]), flush=True)
    dg.close()
