"""GPU parity of batched personalised PageRank (nkpr_ppr_batch) against the oracle.

Each column must equal the oracle's PageRank loop run with that column's teleport vector:
  FP64, reference stopping rule: the same per-column iteration count and an L1 difference
      <= 1e-12 per column (summation order only).
  MIXED: per-column L1 error <= tol (1e-8) against the sparse direct solve, every column
      certified (error bound <= tol).
  FP32 with the certified stop at 1e-6: L1 error <= 1e-6 + 2e-6 (FP32 rounding of the sum).
"""
import numpy as np
import pytest

import oracle
from graphs import random_graph, skewed_graph
from networkit_amd import capi
from networkit_amd import generators as gen

pytestmark = pytest.mark.gpu


def device_graph(g):
    w32 = None if g.in_w is None else g.in_w.astype(np.float32)
    return capi.DeviceGraph(g.z, g.in_ptr, g.in_src.astype(np.int32), w32, exists=g.exists,
                            weight64=g.in_w)


def teleport(z, seeds, weights=None):
    t = np.zeros(z)
    np.add.at(t, np.asarray(seeds, np.int64), 1.0 if weights is None else np.asarray(weights))
    return t / t.sum()


def random_seeds(z, k, rng, max_size=50, exists=None):
    pool = np.arange(z) if exists is None else np.nonzero(exists)[0]
    return [rng.choice(pool, int(rng.integers(1, max_size + 1)), replace=True) for _ in range(k)]


def check_fp64(g, seeds, weights=None, norm=capi.NKPR_L2, sinks=0, tol=1e-8, **kw):
    dg = device_graph(g)
    p = capi.default_params(norm=norm, sink_handling=sinks, tol=tol, **kw)
    got, st, its, _ = dg.ppr_batch(seeds, seed_weights=weights, params=p)
    for j, s in enumerate(seeds):
        t = teleport(g.z, s, None if weights is None else weights[j])
        ref, it, _ = oracle.pagerank(g, tol=tol, norm=norm, sink_mode=sinks, teleport=t)
        assert its[j] == it, (j, its[j], it)
        assert np.abs(got[j] - ref).sum() <= 1e-12, j
    return st


def test_batch_fp64_matches_oracle_weighted_directed():
    z, u, v, w = skewed_graph(3000, 40000, 31)
    g = oracle.in_lists(z, u, v, w, True)
    rng = np.random.default_rng(1)
    st = check_fp64(g, random_seeds(z, 16, rng))
    assert st["sweeps_fp64"] >= st["max_iterations_col"]


def test_batch_fp64_l1_distribute_sinks_weights():
    z, u, v, w = random_graph(2500, 6000, 32, directed=True, weighted=True)
    g = oracle.in_lists(z, u, v, w, True)
    rng = np.random.default_rng(2)
    seeds = random_seeds(z, 7, rng)
    weights = [rng.uniform(0.1, 3.0, len(s)) for s in seeds]
    check_fp64(g, seeds, weights, norm=capi.NKPR_L1, sinks=capi.NKPR_SINK_DISTRIBUTE)


def test_batch_single_column_and_unweighted_undirected():
    z, u, v, w = random_graph(1500, 5000, 33, directed=False)
    g = oracle.in_lists(z, u, v, None, False)
    check_fp64(g, [[3]])
    check_fp64(g, random_seeds(z, 5, np.random.default_rng(3)), sinks=1)


def test_batch_uniform_column_equals_run():
    # a column whose teleport is uniform over all nodes is plain PageRank
    z, u, v, w = skewed_graph(4000, 30000, 34)
    g = oracle.in_lists(z, u, v, w, True)
    dg = device_graph(g)
    p = capi.default_params()
    got, _, its, _ = dg.ppr_batch([np.arange(z), [0, 1]], params=p)
    x = np.full(z, 1.0 / z)
    st = dg.run(x, capi.default_params())
    assert its[0] == st["iterations"]
    assert np.abs(got[0] - x).sum() <= 1e-12


def test_batch_heavy_rows_split():
    # in-hubs above every autotuner threshold: rows are cut into pieces and recombined
    rng = np.random.default_rng(35)
    z = 6000
    u = rng.integers(0, z, 60000)
    v = rng.integers(0, z, 60000)
    v[:12000] = 7          # one row with 12000 in-edges
    v[12000:17000] = 11    # and one with 5000
    g = oracle.in_lists(z, u, v, rng.uniform(0.5, 1.5, len(u)), True)
    check_fp64(g, random_seeds(z, 9, rng))


def test_batch_deleted_slots():
    z, u, v, w = random_graph(2000, 8000, 36, directed=True)
    exists = np.ones(z, np.uint8)
    dead = [5, 17, 999]
    exists[dead] = 0
    keep = ~np.isin(u, dead) & ~np.isin(v, dead)
    g = oracle.in_lists(z, u[keep], v[keep], None, True, exists=exists)
    rng = np.random.default_rng(4)
    check_fp64(g, random_seeds(z, 6, rng, exists=exists))


@pytest.fixture
def sliced(monkeypatch):
    monkeypatch.setenv("NKPR_SLICES", "8")
    monkeypatch.setenv("NKPR_DIRECT_MAX_DEG", "2")


def test_batch_on_sliced_layout(sliced):
    # the batched sweep reads the plain copy of a sliced edge layout
    z, u, v, w = skewed_graph(20000, 200000, 37)
    g = oracle.in_lists(z, u, v, w, True)
    check_fp64(g, random_seeds(z, 12, np.random.default_rng(5)))


def converged_columns(g, seeds, sink_mode=0):
    """The oracle run to an L1 step of 1e-15 (certified error ~1e-14) per column."""
    out = []
    for s in seeds:
        r, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000,
                                  sink_mode=sink_mode, teleport=teleport(g.z, s))
        out.append(r)
    return out


@pytest.mark.parametrize("sink_mode", [0, 1])
def test_batch_mixed_certified(sink_mode):
    u, v = gen.barabasi_albert(20000, 8, seed=4)
    g = oracle.in_lists(20000, u, v, None, False)
    rng = np.random.default_rng(6)
    seeds = random_seeds(g.z, 24, rng, max_size=1000)
    dg = device_graph(g)
    p = capi.default_params(precision=capi.NKPR_MIXED, tol=1e-8, sink_handling=sink_mode)
    got, st, its, bound = dg.ppr_batch(seeds, params=p)
    assert (bound <= 1e-8).all()
    assert st["refinements"] >= 1 and st["sweeps_fp64"] >= 2
    ref = converged_columns(g, seeds, sink_mode)
    for j in range(len(seeds)):
        assert np.abs(got[j] - ref[j]).sum() <= 1e-8, j


def test_batch_mixed_weighted_directed_with_sinks():
    z, u, v, w = skewed_graph(8000, 60000, 38)
    g = oracle.in_lists(z, u, v, w, True)
    rng = np.random.default_rng(7)
    seeds = random_seeds(z, 10, rng)
    dg = device_graph(g)
    for sink_mode in (0, 1):
        p = capi.default_params(precision=capi.NKPR_MIXED, tol=1e-9, sink_handling=sink_mode)
        got, st, its, bound = dg.ppr_batch(seeds, params=p)
        assert (bound <= 1e-9).all()
        ref = converged_columns(g, seeds, sink_mode)
        for j in range(len(seeds)):
            assert np.abs(got[j] - ref[j]).sum() <= 1e-9, j


def test_batch_fp32_certified_stop():
    z, u, v, w = skewed_graph(5000, 50000, 39)
    g = oracle.in_lists(z, u, v, w, True)
    seeds = random_seeds(z, 8, np.random.default_rng(8))
    dg = device_graph(g)
    p = capi.default_params(precision=capi.NKPR_FP32, stop=capi.NKPR_STOP_L1_ERROR, tol=1e-6)
    got, st, its, bound = dg.ppr_batch(seeds, params=p)
    ref = converged_columns(g, seeds)
    for j in range(len(seeds)):
        assert np.abs(got[j] - ref[j]).sum() <= 1e-6 + 2e-6, j


def test_batch_device_scores():
    torch = pytest.importorskip("torch")
    z, u, v, w = random_graph(3000, 12000, 40, directed=True, weighted=True)
    g = oracle.in_lists(z, u, v, w, True)
    seeds = random_seeds(z, 5, np.random.default_rng(9))
    dg = device_graph(g)
    host, _, _, _ = dg.ppr_batch(seeds, params=capi.default_params())
    x = torch.full((5, z), 1.0 / z, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    dg.ppr_batch(seeds, scores=x, params=capi.default_params(scores_on_device=1))
    assert np.abs(x.cpu().numpy() - host).max() == 0.0


def test_batch_invalid_inputs():
    z, u, v, w = random_graph(100, 300, 41, directed=True)
    g = oracle.in_lists(z, u, v, None, True)
    dg = device_graph(g)
    with pytest.raises(capi.NkprError) as e:
        dg.ppr_batch([[1, 2], [100]])
    assert e.value.code == capi.NKPR_ERR_INVALID
    with pytest.raises(capi.NkprError):
        dg.ppr_batch([[1, 2], []])
    with pytest.raises(capi.NkprError):
        dg.ppr_batch([[1]], params=capi.default_params(normalized=1))
    with pytest.raises(capi.NkprError):
        dg.ppr_batch([[1, 2]], seed_weights=[[0.0, 0.0]])
    # the handle still works after rejected calls
    got, _, _, _ = dg.ppr_batch([[1, 2]])
    assert abs(got[0].sum() - 1.0) < 1e-12
