"""The default paths of large graphs against the oracle, at the sizes where they switch on.

A graph (or a batch union) with 2^26 or more in-list entries takes both binned copies by
default: the GMRES corrections of a mixed solve run the binned FP32 operator (pass-1 places
derived from group data) and every FP64 sweep runs the binned FP64 sweep (with 4-byte weights
when every weight is an FP32 number, as R-MAT's are).  Each case is checked against the oracle
run on the same in-lists, not against another mode of the library:
  FP64 with the reference stopping rule (L2 step <= 1e-8): the oracle's iteration count and an
      L1 difference <= 1e-12 (summation order only);
  MIXED with a certified L1 error <= 1e-9: L1 <= 1e-9 against the oracle run on to an L1 step
      of 1e-12, and the same top-1000 ranking (ties within 1e-12 excepted).
The oracle's converged run starts from its own reference-rule result, rescaled to the sum the
unnormalised fixed point has (below), which saves the sweeps the two runs share.
"""
import os

import numpy as np
import pytest

import oracle
from networkit_amd import capi
from networkit_amd import generators as gen

pytestmark = pytest.mark.gpu

DAMP = 0.85
BINNED_ENV = ("NKPR_BINNED", "NKPR_BINNED64", "NKPR_BINNED_ITEM", "NKPR_BINNED_POS",
              "NKPR_COMPACT64", "NKPR_CLOSED_START")


@pytest.fixture(autouse=True)
def default_paths(monkeypatch):
    for k in BINNED_ENV:
        monkeypatch.delenv(k, raising=False)
    monkeypatch.setenv("NKPR_DEBUG", "1")


def converged_start(g, y, sink_mode):
    """A start vector for the oracle's converged run from its normalised reference-rule result
    y.  With distributed sinks the fixed point sums to 1.  Without, rank reaching a sink is lost
    each sweep: the fixed point x satisfies sum(x) = d (sum(x) - S) + (1 - d), S the sink mass,
    so with S = sigma sum(x) (sigma = sink share of y), sum(x) = (1 - d) / (1 - d + d sigma)."""
    if sink_mode == 1:
        return y
    _, cnt = oracle.out_degrees(g)
    ex = np.ones(g.z, bool) if g.exists is None else g.exists.astype(bool)
    sigma = float(y[(cnt == 0) & ex].sum())
    return y * ((1.0 - DAMP) / (1.0 - DAMP + DAMP * sigma))


def oracle_pair(g, sink_mode=0):
    """(reference-rule scores, iterations, converged scores) of the oracle."""
    step, it, _ = oracle.pagerank(g, sink_mode=sink_mode)
    print("oracle: reference rule %d sweeps" % it, flush=True)
    tight, it2, _ = oracle.pagerank(g, tol=1e-12, norm=oracle.L1, max_iterations=20000,
                                    sink_mode=sink_mode, start=converged_start(g, step, sink_mode))
    print("oracle: converged run %d more sweeps" % it2, flush=True)
    return step, it, tight


def assert_same_top(ref, got, top=1000, ties=1e-12):
    order = np.lexsort((np.arange(len(ref)), -ref))[:top]
    got_order = np.lexsort((np.arange(len(got)), -got))[:top]
    if np.array_equal(order, got_order):
        return
    # differences only among scores within `ties` of each other
    for a, b in zip(order, got_order):
        if a != b:
            assert abs(ref[a] - ref[b]) <= ties, (a, b, ref[a], ref[b])


def run_device(dg, z, precision, tol, stop, sink_mode=0):
    import torch
    x = torch.full((z,), 1.0 / z, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = dg.run(x, capi.default_params(scores_on_device=1, precision=precision, tol=tol,
                                       stop=stop, sink_handling=sink_mode))
    torch.cuda.synchronize()
    return x.cpu().numpy(), st


def host_lists(g):
    colptr = g["colptr"].cpu().numpy().astype(np.uint64)
    src = g["src"].cpu().numpy().astype(np.uint64)
    w = None if g["weight"] is None else g["weight"].cpu().numpy().astype(np.float64)
    return oracle.InLists(g["z"], colptr, src, w)


def check_single(dg, og, capfd, sink_mode=0):
    z = og.z
    step, it, tight = oracle_pair(og, sink_mode)
    x64, st64 = run_device(dg, z, capi.NKPR_FP64, 1e-8, capi.NKPR_STOP_STEP, sink_mode)
    err = capfd.readouterr().err
    assert "binned FP64 layout" in err, err[-2000:]
    print("FP64: %d sweeps (oracle %d), L1 %.3e" % (st64["iterations"], it,
                                                    np.abs(x64 - step).sum()), flush=True)
    assert st64["iterations"] == it
    assert np.abs(x64 - step).sum() <= 1e-12
    xm, stm = run_device(dg, z, capi.NKPR_MIXED, 1e-9, capi.NKPR_STOP_L1_ERROR, sink_mode)
    err = capfd.readouterr().err
    assert "binned FP32 layout" in err, err[-2000:]
    e = np.abs(xm - tight).sum()
    print("MIXED: %d + %d sweeps, bound %.3e, L1 %.3e" % (stm["sweeps_fp32"], stm["sweeps_fp64"],
                                                         stm["l1_error_bound"], e), flush=True)
    assert stm["l1_error_bound"] <= 1e-9
    assert e <= 1e-9
    assert_same_top(tight, xm)


def test_rmat23_default_paths(capfd):
    """R-MAT scale 23, weighted (FP32 weights, so the FP64 copy keeps 4-byte weights)."""
    import torch
    g = gen.rmat(23, 16, seed=4, device="cuda")
    assert g["src"].numel() >= 1 << 26
    dg = capi.DeviceGraph(g["z"], g["colptr"], g["src"], g["weight"], device=True)
    og = host_lists(g)
    del g
    check_single(dg, og, capfd)
    dg.close()
    torch.cuda.empty_cache()


def test_rhg_default_paths(capfd):
    """Undirected threshold random hyperbolic graph, n = 4M, average degree 20 (unit weights;
    the slowly mixing, skewed-degree case)."""
    import torch
    g = gen.rhg(4_000_000, 20.0, 3.0, seed=3)
    assert g["src"].numel() >= 1 << 26
    dg = capi.DeviceGraph(g["z"], g["colptr"], g["src"], None, device=True)
    og = host_lists(g)
    del g
    check_single(dg, og, capfd, sink_mode=1)
    dg.close()
    torch.cuda.empty_cache()


def test_graph_batch_union_default_paths(capfd):
    """40 undirected Erdos-Renyi graphs, n ~ U{100k..250k}, average degree 12, in one batch
    handle (about 8.4e7 arcs): every graph against its own oracle run."""
    import torch
    rng = np.random.default_rng(12)
    G = 40
    sizes = rng.integers(100_000, 250_001, size=G)
    graphs, ptrs, srcs, offs = [], [np.zeros(1, np.uint64)], [], [0]
    base = 0
    for i in range(G):
        n = int(sizes[i])
        u, v = gen.erdos_renyi(n, 12.0 / (n - 1), seed=int(rng.integers(1, 1 << 62)))
        og = oracle.in_lists(n, u, v, None, False)
        graphs.append(og)
        ptrs.append(og.in_ptr[1:] + np.uint64(base))
        srcs.append(og.in_src.astype(np.int32) + np.int32(offs[-1]))
        base += og.m
        offs.append(offs[-1] + n)
    colptr = np.concatenate(ptrs)
    src = np.concatenate(srcs)
    assert base >= 1 << 26
    dg = capi.DeviceGraph(offs[-1], colptr, src, graph_offsets=np.array(offs, np.uint64))
    del colptr, src
    for precision, tol, stop in ((capi.NKPR_FP64, 1e-8, capi.NKPR_STOP_STEP),
                                 (capi.NKPR_MIXED, 1e-9, capi.NKPR_STOP_L1_ERROR)):
        x = np.concatenate([np.full(gr.z, 1.0 / gr.z) for gr in graphs])
        st, its, _ = dg.run_batch(x, precision=precision, tol=tol, stop=stop)
        err = capfd.readouterr().err
        assert ("binned FP64 layout" if precision == capi.NKPR_FP64 else "binned FP32 layout") in err
        if precision == capi.NKPR_MIXED:
            assert st["l1_error_bound"] <= 1e-9
        worst = 0.0
        for q, og in enumerate(graphs):
            if q % 10 == 0:
                print("graph %d of %d against the oracle" % (q, G), flush=True)
            xq = x[offs[q]:offs[q + 1]]
            if precision == capi.NKPR_FP64:
                ref, it, _ = oracle.pagerank(og)
                assert its[q] == it, (q, its[q], it)
                e = np.abs(xq - ref).sum()
                assert e <= 1e-12, (q, e)
            else:
                step, _, _ = oracle.pagerank(og)
                ref, _, _ = oracle.pagerank(og, tol=1e-12, norm=oracle.L1, max_iterations=20000,
                                            start=converged_start(og, step, 0))
                e = np.abs(xq - ref).sum()
                assert e <= 1e-9, (q, e)
                assert_same_top(ref, xq)
            worst = max(worst, e)
        print("%s: worst per-graph L1 %.3e" % ("FP64" if precision == capi.NKPR_FP64 else "MIXED",
                                               worst), flush=True)
    dg.close()
    torch.cuda.empty_cache()
