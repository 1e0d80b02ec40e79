"""GPU parity of the PageRank drop-in against the oracle and the reference's known answers.

Bars (written into each assertion):
  FP64 mode, reference stopping rule: the same iteration count as the oracle and an L1
  difference <= 1e-12 (rounding of a different summation order only); the same ranking.
  Mixed mode: L1 error <= tol (1e-8) against a tightly converged FP64 oracle and the same
  top-1000 ranking up to ties within 1e-12.
"""
import ctypes

import numpy as np
import pytest

import networkit_amd as nk
import oracle
from graphs import (EIGHT_PR, EIGHT_PR_NORMALIZED, WIKI_TOP, WIKI_TOP_NORMALIZED, eight_node,
                    random_graph, skewed_graph, wiki_vote)
from networkit_amd import capi
from networkit_amd import generators as gen

pytestmark = pytest.mark.gpu

NORMS = {0: nk.centrality.Norm.l1norm, 1: nk.centrality.Norm.l2norm}


def build_graph(z, u, v, w, directed, remove=()):
    G = nk.Graph(z, w is not None, directed)
    G.addEdges(u, v, w)
    for x in remove:
        G.removeNode(x)
    return G


def oracle_graph(z, u, v, w, directed, remove=()):
    keep = np.ones(len(u), bool)
    for x in remove:
        keep &= (u != x) & (v != x)
    exists = None
    if remove:
        exists = np.ones(z, np.uint8)
        exists[list(remove)] = 0
    return oracle.in_lists(z, u[keep], v[keep], None if w is None else w[keep], directed,
                           exists=exists)


def run_api(G, damp=0.85, tol=1e-8, norm=1, normalized=False, sinks=0, max_it=None,
            precision=0, pr=None):
    if pr is None:
        pr = nk.centrality.PageRank(G, damp, tol, normalized, sinks)
    pr.norm = NORMS[norm]
    if max_it is not None:
        pr.maxIterations = max_it
    pr.precision = precision
    pr.run()
    return pr, np.array(pr.scores())


def assert_same_ranking(a, b, exists=None, ties=1e-12, top=None):
    ra = oracle.ranking(a, exists)
    rb = oracle.ranking(b, exists)
    if top is not None:
        ra, rb = ra[:top], rb[:top]
    diff = np.nonzero(ra != rb)[0]
    for i in diff:
        # positions may swap only between nodes whose scores tie within `ties`
        assert abs(a[ra[i]] - a[rb[i]]) <= ties, (i, ra[i], rb[i])


def check_fp64_parity(z, u, v, w, directed, remove=(), **kw):
    G = build_graph(z, u, v, w, directed, remove)
    g = oracle_graph(z, u, v, w, directed, remove)
    pr, got = run_api(G, **kw)
    ref, it, mx = oracle.pagerank(g, kw.get("damp", 0.85), kw.get("tol", 1e-8),
                                  kw.get("norm", 1), kw.get("max_it", 2 ** 64 - 1),
                                  kw.get("normalized", False), kw.get("sinks", 0))
    assert pr.numberOfIterations() == it
    scale = 1.0 if not kw.get("normalized") else max(1.0, np.abs(ref).max())
    assert np.abs(got - ref).sum() <= 1e-12 * scale * max(1, z / 1000)
    assert pr.maximum() == pytest.approx(mx, rel=1e-12)
    assert_same_ranking(ref, got, g.exists)
    return pr, got, ref


# ---------------------------------------------------------------- reference known answers
@pytest.mark.parametrize("normalized", [False, True])
def test_eight_node_known_answers(normalized):
    z, u, v, w = eight_node()
    pr, got, ref = check_fp64_parity(z, u, v, w, False, normalized=normalized)
    expect = EIGHT_PR_NORMALIZED if normalized else EIGHT_PR
    assert np.allclose(got, expect, atol=1e-4)


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("norm", [0, 1])
@pytest.mark.parametrize("normalized", [False, True])
def test_wiki_vote_known_answers(directed, norm, normalized):
    z, u, v = wiki_vote(directed)
    pr, got, ref = check_fp64_parity(z, u, v, None, directed, norm=norm,
                                     normalized=normalized)
    r = pr.ranking()
    top, val = (WIKI_TOP_NORMALIZED if normalized else WIKI_TOP)[directed]
    assert r[0][0] == top
    assert abs(r[0][1] - val) <= (1e-4 if normalized else 1e-3)
    # re-run with maxIterations = 2 starts from the previous scores (PageRank.cpp:26)
    pr.maxIterations = 2
    pr.run()
    assert pr.numberOfIterations() <= 2
    g = oracle_graph(z, u, v, None, directed)
    ref2, it2, _ = oracle.pagerank(g, norm=norm, max_iterations=2, normalized=normalized,
                                   start=got)
    assert pr.numberOfIterations() == it2
    assert np.abs(np.array(pr.scores()) - ref2).sum() <= 1e-12 * max(1, np.abs(ref2).max())


# ---------------------------------------------------------------- oracle parity, FP64 mode
def test_config1_er_undirected_unweighted():
    u, v = gen.erdos_renyi(10000, 1e-3, seed=1)
    for sinks in (0, 1):
        check_fp64_parity(10000, u, v, None, False, sinks=sinks)


@pytest.mark.parametrize("sinks", [0, 1])
@pytest.mark.parametrize("normalized", [False, True])
def test_rmat_weighted_directed(sinks, normalized):
    g = gen.rmat(13, 8, seed=2)
    z = g["z"]
    rows = np.repeat(np.arange(z), np.diff(g["colptr"].astype(np.int64)))
    check_fp64_parity(z, g["src"], rows, g["weight"].astype(np.float64), True, sinks=sinks,
                      normalized=normalized)


def test_heavy_rows_split_and_vector_items():
    # one row with 40000 in-edges (split across blocks), one with 5000 (one block), the rest
    # a sparse random graph with self-loops and multi-edges
    rng = np.random.default_rng(3)
    z = 50000
    u = np.concatenate([np.arange(1, 40001), rng.integers(0, z, 5000), rng.integers(0, z, 60000)])
    v = np.concatenate([np.zeros(40000, np.int64), np.full(5000, 7), rng.integers(0, z, 60000)])
    w = rng.uniform(0.5, 1.5, len(u))
    for directed in (True, False):
        check_fp64_parity(z, u, v, w, directed, norm=0)
        check_fp64_parity(z, u, v, None, directed, sinks=1, normalized=True)


def test_skewed_directed_weighted():
    z, u, v, w = skewed_graph(30000, 400000, 5)
    check_fp64_parity(z, u, v, w, True)


@pytest.fixture
def sliced(monkeypatch):
    """Force the source-sliced schedule (8 slices, rows with > 2 in-edges sliced) on small
    graphs; by default it is used only from 2^20 nodes up."""
    monkeypatch.setenv("NKPR_SLICES", "8")
    monkeypatch.setenv("NKPR_DIRECT_MAX_DEG", "2")


@pytest.mark.parametrize("normalized", [False, True])
def test_sliced_schedule_parity(sliced, normalized):
    z, u, v, w = skewed_graph(30000, 400000, 5)
    check_fp64_parity(z, u, v, w, True, normalized=normalized)
    check_fp64_parity(z, u, v, None, False, sinks=1, norm=0)
    # heavy rows: split chunks and vector segments inside slices
    rng = np.random.default_rng(3)
    z = 200000
    u = np.concatenate([np.arange(1, 150001), rng.integers(0, z, 20000), rng.integers(0, z, 300000)])
    v = np.concatenate([np.zeros(150000, np.int64), np.full(20000, 7), rng.integers(0, z, 300000)])
    w = rng.uniform(0.5, 1.5, len(u))
    check_fp64_parity(z, u, v, w, True, norm=0)


def test_sliced_mixed_precision(sliced):
    g = gen.rmat(15, 16, seed=2)
    z = g["z"]
    rows = np.repeat(np.arange(z), np.diff(g["colptr"].astype(np.int64)))
    w = g["weight"].astype(np.float64)
    G = build_graph(z, g["src"], rows, w, True)
    ref, _, _ = oracle.pagerank(oracle_graph(z, g["src"], rows, w, True), tol=1e-15, norm=0,
                                max_iterations=3000)
    pr = nk.centrality.PageRank(G, 0.85, 1e-8)
    pr.precision = nk.centrality.Precision.Mixed
    pr.run()
    got = np.array(pr.scores())
    assert np.abs(got - ref).sum() <= 1e-8
    assert_same_ranking(ref, got, top=1000)


def test_deleted_nodes_self_loops_multi_edges():
    z, u, v, w = random_graph(3000, 12000, 9, directed=True, weighted=True)
    for directed in (True, False):
        pr, got, ref = check_fp64_parity(z, u, v, w, directed, remove=(0, 17, 2999),
                                         normalized=True)
        assert got[17] == 1.0 / 2997  # deleted slots keep the start value


def test_max_iterations_zero_and_one():
    z, u, v, w = random_graph(500, 2000, 2)
    check_fp64_parity(z, u, v, w, True, max_it=0)
    check_fp64_parity(z, u, v, w, True, max_it=1)


def test_edgeless_and_tiny_graphs():
    G = nk.Graph(5, False, True)
    pr, got = run_api(G)
    assert np.allclose(got, 0.2, rtol=0, atol=1e-16)
    # sweep 1 moves every score to (1-d)/n, sweep 2 repeats it: the step is 0
    _, it, _ = oracle.pagerank(oracle.in_lists(5, [], [], None, True))
    assert pr.numberOfIterations() == it == 2
    G = nk.Graph(1, False, False)
    G.addEdge(0, 0)
    pr, got = run_api(G)
    assert got.tolist() == [1.0]


def test_centralization_and_score():
    z, u, v, w = random_graph(400, 1600, 4)
    G = build_graph(z, u, v, None, True)
    pr, got = run_api(G, normalized=True)
    assert pr.score(5) == got[5]
    c = pr.centralization()
    mx = pr.maximum()
    center = got.max()
    assert c == pytest.approx((center - got).sum() / (mx - got).sum())


def test_graph_change_reexports():
    z, u, v, w = random_graph(300, 900, 6)
    G = build_graph(z, u, v, None, True)
    pr, a = run_api(G)
    G.addEdge(0, 1)
    pr2, b = run_api(G)
    g = oracle.in_lists(z, np.append(u, 0), np.append(v, 1), None, True)
    ref, _, _ = oracle.pagerank(g)
    assert np.abs(b - ref).sum() <= 1e-12


# ---------------------------------------------------------------- mixed precision
@pytest.mark.parametrize("sinks", [0, 1])
def test_mixed_precision_reaches_tol(sinks):
    g = gen.rmat(15, 16, seed=2)
    z = g["z"]
    rows = np.repeat(np.arange(z), np.diff(g["colptr"].astype(np.int64)))
    w = g["weight"].astype(np.float64)
    G = build_graph(z, g["src"], rows, w, True)
    og = oracle_graph(z, g["src"], rows, w, True)
    ref, _, _ = oracle.pagerank(og, tol=1e-15, norm=0, max_iterations=3000, sink_mode=sinks)
    pr = nk.centrality.PageRank(G, 0.85, 1e-8, False, sinks)
    pr.precision = nk.centrality.Precision.Mixed
    pr.run()
    got = np.array(pr.scores())
    st = pr.stats()
    err = np.abs(got - ref).sum()
    print("mixed: L1 %.3e bound %.3e sweeps fp32 %d fp64 %d" % (
        err, st["l1_error_bound"], st["sweeps_fp32"], st["sweeps_fp64"]))
    assert err <= 1e-8
    assert st["l1_error_bound"] <= 1e-8 and err <= st["l1_error_bound"]
    assert st["sweeps_fp64"] <= 4 and st["sweeps_fp32"] > 0
    assert pr.numberOfIterations() == st["sweeps_fp32"] + st["sweeps_fp64"]
    assert_same_ranking(ref, got, top=1000)


def test_fp32_mode_is_close():
    z, u, v, w = random_graph(5000, 40000, 12, weighted=True)
    G = build_graph(z, u, v, w, True)
    pr = nk.centrality.PageRank(G, 0.85, 1e-5)
    pr.precision = nk.centrality.Precision.FP32
    pr.certifiedStop = True
    pr.run()
    got = np.array(pr.scores())
    ref, _, _ = oracle.pagerank(oracle_graph(z, u, v, w, True), tol=1e-14, norm=0)
    # certified bound 1e-5 on the FP32 iterate; FP32 rounding adds ~1e-6 at this size
    assert np.abs(got - ref).sum() < 1e-5 + 2e-6
    assert pr.stats()["sweeps_fp64"] == 0
    # FP32 iterates certify nothing: the bound is reported as +inf
    assert pr.stats()["l1_error_bound"] == float("inf")


# ---------------------------------------------------------------- the C ABI directly
def test_capi_device_pointers_and_device_scores():
    import torch
    g = gen.rmat(14, 16, seed=5, device="cuda")
    dg = capi.DeviceGraph(g["z"], g["colptr"], g["src"], g["weight"], device=True)
    info = dg.info()
    assert info["n"] == g["z"] and info["m"] == g["src"].numel()
    x = torch.full((g["z"],), 1.0 / g["z"], dtype=torch.float64, device="cuda")
    st = dg.run(x, capi.default_params(scores_on_device=1, tol=1e-10))
    torch.cuda.synchronize()
    got = x.cpu().numpy()
    h = gen.rmat(14, 16, seed=5, device="cpu")
    assert np.array_equal(h["src"], g["src"].cpu().numpy())
    z = h["z"]
    rows = np.repeat(np.arange(z), np.diff(h["colptr"].astype(np.int64)))
    ref, it, _ = oracle.pagerank(oracle.in_lists(z, h["src"], rows,
                                                 h["weight"].astype(np.float64), True),
                                 tol=1e-10)
    assert st["iterations"] == it
    assert np.abs(got - ref).sum() <= 1e-12


def test_capi_interrupt_flag():
    z, u, v, w = random_graph(1000, 5000, 1)
    c = gen.csc_from_edges(z, u, v, None, True)
    dg = capi.DeviceGraph(z, c["colptr"], c["src"])
    flag = ctypes.c_int32(1)
    p = capi.default_params(interrupt=ctypes.cast(ctypes.pointer(flag), ctypes.c_void_p))
    x = np.full(z, 1.0 / z)
    with pytest.raises(capi.NkprError) as e:
        dg.run(x, p)
    assert e.value.code == capi.NKPR_ERR_INTERRUPTED
    assert np.all(x == 1.0 / z)  # untouched


def slow_graph():
    """A graph and damping factor whose run takes far longer than the interrupt below needs:
    with d = 0.99999 the step shrinks by about 1e-5 per sweep."""
    z, u, v, w = random_graph(20000, 120000, 3, directed=True, weighted=True)
    return z, u, v, w, 0.99999


def test_capi_interrupt_during_a_run():
    """The flag set by another host thread while nkpr_run sweeps: the run stops at its next
    check with NKPR_ERR_INTERRUPTED and the caller's scores are untouched."""
    import threading
    import time
    z, u, v, w, damp = slow_graph()
    c = gen.csc_from_edges(z, u, v, w, True)
    dg = capi.DeviceGraph(z, c["colptr"], c["src"], c["weight"])
    flag = ctypes.c_int32(0)
    p = capi.default_params(damp=damp, tol=0.0, max_iterations=10 ** 8,
                            interrupt=ctypes.cast(ctypes.pointer(flag), ctypes.c_void_p))
    start = np.random.default_rng(4).random(z)
    x = start.copy()
    timer = threading.Timer(0.5, lambda: setattr(flag, "value", 1))
    t0 = time.time()
    timer.start()
    with pytest.raises(capi.NkprError) as e:
        dg.run(x, p)
    timer.join()
    assert e.value.code == capi.NKPR_ERR_INTERRUPTED
    assert time.time() - t0 < 30.0
    assert np.array_equal(x, start)
    # the handle runs normally afterwards (from the same, untouched, start vector)
    flag.value = 0
    st = dg.run(x, capi.default_params())
    ref, it, _ = oracle.pagerank(oracle.in_lists(z, u, v, w, True), start=start)
    assert st["iterations"] == it and np.abs(x - ref).sum() <= 1e-12


def test_sigint_during_run_raises():
    """SIGINT while PageRank.run() sweeps: the drop-in's signal handler (installed for the
    duration of run(), like Aux::SignalHandler) turns it into InterruptException, which reaches
    Python as RuntimeError; a later run() works and Python's own handler is back in place."""
    import signal
    import threading
    import time
    z, u, v, w, damp = slow_graph()
    G = build_graph(z, u, v, w, True)
    pr = nk.centrality.PageRank(G, damp, 0.0)
    pr.maxIterations = 10 ** 8
    got = []
    prev = signal.signal(signal.SIGINT, lambda s, f: got.append(s))
    main = threading.get_ident()
    done = threading.Event()

    def poke():
        # a signal that lands before run() installs its handler only reaches `got`
        while not done.wait(0.25):
            signal.pthread_kill(main, signal.SIGINT)

    th = threading.Thread(target=poke)
    try:
        t0 = time.time()
        th.start()
        with pytest.raises(RuntimeError, match="SIGINT"):
            pr.run()
        done.set()
        th.join()
        assert time.time() - t0 < 30.0
        assert not pr.hasFinished()
        # Python's handler receives SIGINT again once run() has returned
        n0 = len(got)
        signal.pthread_kill(main, signal.SIGINT)
        for _ in range(100):
            if len(got) > n0:
                break
            time.sleep(0.01)
        assert len(got) > n0
        pr2 = nk.centrality.PageRank(G)
        pr2.run()
        ref, it, _ = oracle.pagerank(oracle_graph(z, u, v, w, True))
        assert pr2.numberOfIterations() == it
        assert np.abs(np.array(pr2.scores()) - ref).sum() <= 1e-12
    finally:
        done.set()
        th.join()
        signal.signal(signal.SIGINT, prev)


def test_capi_rejects_bad_input():
    with pytest.raises(capi.NkprError) as e:
        capi.DeviceGraph(3, np.array([0, 1, 2, 3]), np.array([1, 5, 0]))
    assert e.value.code == capi.NKPR_ERR_INVALID
    with pytest.raises(capi.NkprError):
        capi.DeviceGraph(3, np.array([0, 2, 1, 3]), np.array([1, 2, 0]))
    dg = capi.DeviceGraph(3, np.array([0, 1, 2, 3]), np.array([1, 2, 0]))
    with pytest.raises(capi.NkprError):
        dg.run(np.full(3, 1 / 3), capi.default_params(damp=1.5))


@pytest.mark.parametrize("weighted", [True, False])
@pytest.mark.parametrize("sinks", [0, 1])
@pytest.mark.parametrize("item", ["0", "3000"])
def test_mixed_binned_operator(monkeypatch, weighted, sinks, item):
    """The GMRES correction solves through the two-pass binned operator (forced on a graph
    spanning several source chunks and destination bins; item=3000 splits bins over work items,
    whose accumulators are added in order by the reduce kernel): every result still certifies
    at tol and matches the converged oracle, and the solve takes the same refinements and
    (within one) the same operator applications as with the pull-sweep operator.  Both
    operators are accurate to FP32; a wrong product (a lost segment, a wrong bin or sink term)
    would slow the correction solves down."""
    monkeypatch.setenv("NKPR_BINNED_ITEM", item)
    z, u, v, w = skewed_graph(120000, 900000, 61, weighted=weighted)
    g = oracle_graph(z, u, v, w, True)
    ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=0, max_iterations=5000, sink_mode=sinks)
    w32 = None if g.in_w is None else g.in_w.astype(np.float32)
    stats = {}
    for binned in ("0", "1"):
        monkeypatch.setenv("NKPR_BINNED", binned)
        dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), w32, weight64=g.in_w)
        x = np.full(z, 1.0 / z)
        st = dg.run(x, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-9,
                                           stop=capi.NKPR_STOP_L1_ERROR, sink_handling=sinks))
        assert st["l1_error_bound"] <= 1e-9
        assert np.abs(x - ref).sum() <= 1e-9
        stats[binned] = st
    # GMRES stops on the L2 norm of its residual while the certificate is in L1, so a
    # refinement that ends close to its target can leave one of the two solves just short of
    # the bound (one more refinement of a few applications); a wrong operator costs many.
    r0, r1 = int(stats["0"]["refinements"]), int(stats["1"]["refinements"])
    assert abs(r1 - r0) <= 1 and max(r0, r1) <= 4
    slack = 1 if r0 == r1 else 8
    assert abs(int(stats["1"]["sweeps_fp32"]) - int(stats["0"]["sweeps_fp32"])) <= slack


@pytest.mark.parametrize("weighted", [True, False])
@pytest.mark.parametrize("binned", ["0", "1"])
@pytest.mark.parametrize("teleport", [False, True])
def test_mixed_krylov_head_without_sinks(monkeypatch, capfd, weighted, binned, teleport):
    """Without sink handling the sinks are zero columns of P, so GMRES runs on the non-sink
    head (sink_begin rounded up to 64 entries) and each sink's correction is rebuilt from the
    products it computed (e_t = r_t + sum_i c_i (B u_i)_t).  Against full-length vectors
    (NKPR_SINK_SPLIT=0): both certify at tol and match the converged oracle, with the same
    refinements (within one); a wrong tail would leave the sinks' residual in place and the
    certificate unmet.  Pull and binned operators (the binned pass 1 reads the sources past
    sink_begin as 0), with and without a teleport vector."""
    monkeypatch.setenv("NKPR_BINNED", binned)
    monkeypatch.setenv("NKPR_BINNED_ITEM", "3000")
    monkeypatch.setenv("NKPR_DEBUG", "1")
    z, u, v, w = skewed_graph(90000, 700000, 67, weighted=weighted)
    g = oracle_graph(z, u, v, w, True)
    t = None
    if teleport:
        t = np.random.default_rng(3).random(z)
        t /= t.sum()
    ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=0, max_iterations=5000, sink_mode=0,
                                teleport=t)
    w32 = None if g.in_w is None else g.in_w.astype(np.float32)
    stats = {}
    for split in ("0", "1"):
        monkeypatch.setenv("NKPR_SINK_SPLIT", split)
        dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), w32, weight64=g.in_w)
        sinks = dg.info()["sinks"]
        x = np.full(z, 1.0 / z)
        kw = {} if t is None else {"teleport": t.ctypes.data}
        capfd.readouterr()
        st = dg.run(x, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-9,
                                           stop=capi.NKPR_STOP_L1_ERROR, sink_handling=0, **kw))
        err = capfd.readouterr().err
        head = [line for line in err.splitlines() if "Krylov vectors" in line]
        want = z if split == "0" else min(z, (z - sinks + 63) // 64 * 64)
        assert head and all(("%d of %d" % (want, z)) in line for line in head), (head, want)
        assert st["l1_error_bound"] <= 1e-9
        assert np.abs(x - ref).sum() <= 1e-9
        stats[split] = st
    print("sinks %d of %d; refinements %d / %d, applications %d / %d" % (
        sinks, z, stats["0"]["refinements"], stats["1"]["refinements"],
        stats["0"]["sweeps_fp32"], stats["1"]["sweeps_fp32"]), flush=True)
    assert sinks > z // 4
    assert abs(int(stats["1"]["refinements"]) - int(stats["0"]["refinements"])) <= 1


@pytest.mark.parametrize("weighted", [True, False])
@pytest.mark.parametrize("item", ["0", "3000"])
def test_binned_derived_places(monkeypatch, capfd, weighted, item):
    """Pass 1 of both binned copies (the FP32 operator and the FP64 sweep) derives every quad's
    bin-major place from per-group data (a segment-start mask and the first segment of each
    group of 64 quads, one offset per segment) instead of reading the stored place.  The
    derivation is checked against the stored places when a layout is built and is in use here;
    every contribution then goes to the same place, so on one graph handle (the same
    coefficients) the mixed solve (FP32 operator, FP64 residual sweeps with 6-byte
    contributions) and the FP64 power method are bit-identical to the solves that read the
    stored places (NKPR_BINNED_POS=1) and to repeats of themselves, and they certify against
    the converged oracle."""
    monkeypatch.setenv("NKPR_BINNED", "1")
    monkeypatch.setenv("NKPR_BINNED64", "1")
    monkeypatch.setenv("NKPR_BINNED_ITEM", item)
    monkeypatch.setenv("NKPR_DEBUG", "1")
    z, u, v, w = skewed_graph(150000, 1200000, 67, weighted=weighted)
    g = oracle_graph(z, u, v, w, True)
    ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=0, max_iterations=5000)
    w32 = None if g.in_w is None else g.in_w.astype(np.float32)
    dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), w32, weight64=g.in_w)
    capfd.readouterr()
    xs, stats = {}, {}
    for pos in ("0", "1", "0"):
        monkeypatch.setenv("NKPR_BINNED_POS", pos)
        for prec, tol in ((capi.NKPR_MIXED, 1e-9), (capi.NKPR_FP64, 1e-8)):
            x = np.full(z, 1.0 / z)
            st = dg.run(x, capi.default_params(precision=prec, tol=tol,
                                               stop=capi.NKPR_STOP_L1_ERROR))
            assert st["l1_error_bound"] <= tol
            assert np.abs(x - ref).sum() <= tol
            key = (pos, prec)
            if key in xs:
                assert np.array_equal(x, xs[key])
            xs[key], stats[key] = x, st
    err = capfd.readouterr().err
    for copy in ("FP32", "FP64"):
        assert ("binned %s places:" % copy) in err
    assert "derived yes" in err and "derived no" not in err
    for prec in (capi.NKPR_MIXED, capi.NKPR_FP64):
        assert np.array_equal(xs[("0", prec)], xs[("1", prec)])
        for key in ("refinements", "sweeps_fp32", "sweeps_fp64"):
            assert stats[("0", prec)][key] == stats[("1", prec)][key]
    if not weighted:
        # a second handle has the same coefficients (weighted out-degrees are added with
        # atomics, in an order that can differ between handles in the last bit)
        dg2 = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), None)
        x = np.full(z, 1.0 / z)
        dg2.run(x, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-9,
                                       stop=capi.NKPR_STOP_L1_ERROR))
        assert np.array_equal(x, xs[("0", capi.NKPR_MIXED)])


@pytest.mark.parametrize("case", ["edgeless", "single_loop", "star_sinks", "deleted", "one_bin"])
def test_mixed_binned_operator_edge_cases(monkeypatch, case):
    """The binned operator forced on degenerate layouts: no edges (no pass-1 items, every bin
    empty), one node with a self-loop, a star whose leaves are all sinks (one source chunk),
    deleted slots, and a graph inside one bin.  Each mixed solve certifies and matches the
    converged oracle (distributed sinks, so the sink term of the operator is exercised)."""
    monkeypatch.setenv("NKPR_BINNED", "1")
    monkeypatch.setenv("NKPR_BINNED_ITEM", "8")
    exists = None
    if case == "edgeless":
        z, u, v, w = 7, np.zeros(0, np.int64), np.zeros(0, np.int64), None
    elif case == "single_loop":
        z, u, v, w = 1, np.array([0]), np.array([0]), None
    elif case == "star_sinks":
        z = 40000
        u = np.zeros(z - 1, np.int64)
        v = np.arange(1, z, dtype=np.int64)
        w = np.random.default_rng(3).uniform(0.5, 1.5, z - 1)
    elif case == "deleted":
        z, u, v, w = random_graph(3000, 12000, 63, weighted=True)
        exists = np.ones(z, np.uint8)
        exists[[5, 77, 2999]] = 0
        keep = (exists[u] == 1) & (exists[v] == 1)
        u, v, w = u[keep], v[keep], w[keep]
    else:
        z, u, v, w = random_graph(5000, 40000, 64, weighted=False)
    g = oracle.in_lists(z, u, v, w, True, exists=exists)
    ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=0, max_iterations=20000, sink_mode=1)
    w32 = None if g.in_w is None else g.in_w.astype(np.float32)
    dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), w32, exists=exists,
                          weight64=g.in_w)
    x = np.full(z, 1.0 / max(int(g.exists.sum()) if g.exists is not None else z, 1))
    st = dg.run(x, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-10,
                                       stop=capi.NKPR_STOP_L1_ERROR, sink_handling=1))
    assert st["l1_error_bound"] <= 1e-10
    ex = np.ones(z, bool) if exists is None else exists.astype(bool)
    assert np.abs(x[ex] - ref[ex]).sum() <= 1e-10


@pytest.mark.parametrize("weights", ["fp64", "fp32", None])
@pytest.mark.parametrize("sinks", [0, 1])
@pytest.mark.parametrize("normalized", [False, True])
@pytest.mark.parametrize("item", ["0", "2000"])
def test_fp64_binned_sweep_parity(monkeypatch, weights, sinks, normalized, item):
    """The reference FP64 iteration with every sweep through the binned FP64 passes (forced on
    a graph of several source chunks and destination bins; item=2000 splits bins over work
    items, finished by the reduce kernel): the oracle's iteration count, L1 difference
    <= 1e-12, maximum() to 1e-12 and the same ranking, with deleted nodes.  Weights that are
    all FP32 numbers take the copy with 4-byte weights (x / deg staged in the table)."""
    monkeypatch.setenv("NKPR_BINNED64", "1")
    monkeypatch.setenv("NKPR_BINNED_ITEM", item)
    z, u, v, w = skewed_graph(50000, 400000, 65, weighted=weights is not None)
    if weights == "fp32":
        w = w.astype(np.float32).astype(np.float64)
    check_fp64_parity(z, u, v, w, True, remove=(3, 4000, 49999), sinks=sinks,
                      normalized=normalized, tol=1e-10)


def test_fp64_binned_sweep_undirected_and_l1(monkeypatch):
    monkeypatch.setenv("NKPR_BINNED64", "1")
    z, u, v, w = random_graph(30000, 150000, 66, directed=False, weighted=True)
    check_fp64_parity(z, u, v, w, False, norm=0, tol=1e-11)


def test_fp64_binned_sweep_teleport_and_mixed(monkeypatch):
    """Sweeps with a per-node right-hand side (teleport vector) and the FP64 residual sweeps of
    the mixed path through the binned FP64 passes."""
    monkeypatch.setenv("NKPR_BINNED64", "1")
    z, u, v, w = skewed_graph(40000, 300000, 67)
    g = oracle.in_lists(z, u, v, w, True)
    rng = np.random.default_rng(8)
    t = np.zeros(z)
    np.add.at(t, rng.choice(z, 50), rng.uniform(0.5, 2.0, 50))
    t /= t.sum()
    dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), g.in_w.astype(np.float32),
                          weight64=g.in_w)
    x = np.full(z, 1.0 / z)
    st = dg.run(x, capi.default_params(teleport=t.ctypes.data, sink_handling=1))
    ref, it, _ = oracle.pagerank(g, sink_mode=1, teleport=t)
    assert st["iterations"] == it
    assert np.abs(x - ref).sum() <= 1e-12
    ref2, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000)
    for binned32 in ("0", "1"):
        monkeypatch.setenv("NKPR_BINNED", binned32)
        dg2 = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), g.in_w.astype(np.float32),
                               weight64=g.in_w)
        x2 = np.full(z, 1.0 / z)
        st2 = dg2.run(x2, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-10,
                                              stop=capi.NKPR_STOP_L1_ERROR))
        assert st2["l1_error_bound"] <= 1e-10
        assert np.abs(x2 - ref2).sum() <= 1e-10


@pytest.mark.parametrize("weights", ["fp32", "fp64", None])
def test_mixed_residual_sweeps_with_6_byte_contributions(monkeypatch, weights):
    """The mixed solve's residual sweeps through the binned FP64 copy store every contribution
    as an FP32 head and a 16-bit tail (within 2^-39 of it).  Against the same solve with 8-byte
    contributions (NKPR_COMPACT64=0): both certify and match the converged oracle at tol, with
    the same refinements (within one), and the compact run's bound carries the slack."""
    monkeypatch.setenv("NKPR_BINNED64", "1")
    z, u, v, w = skewed_graph(50000, 400000, 71, weighted=weights is not None)
    if weights == "fp32":
        w = w.astype(np.float32).astype(np.float64)
    elif weights == "fp64":
        w = w * (1.0 + 1e-9)  # not FP32 numbers: the copy keeps FP64 coefficients
    g = oracle.in_lists(z, u, v, w, True)
    ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("NKPR_COMPACT64", flag)
        w32 = None if g.in_w is None else g.in_w.astype(np.float32)
        dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), w32, weight64=g.in_w)
        x = np.full(z, 1.0 / z)
        st = dg.run(x, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-10,
                                           stop=capi.NKPR_STOP_L1_ERROR))
        assert st["l1_error_bound"] <= 1e-10, flag
        assert np.abs(x - ref).sum() <= 1e-10, flag
        out[flag] = (x, st)
    (xc, sc), (x8, s8) = out["1"], out["0"]
    # a residual differs by at most d 2^-39 ||x||_1 (~1.5e-12 here), below the FP32 rounding
    # that the correction solves add anyway: the solves take the same path up to a refinement
    # that ends close to the bound, and both results are certified within tol of the solution
    assert abs(int(sc["refinements"]) - int(s8["refinements"])) <= 1
    assert np.abs(xc - x8).sum() <= 2e-10
    slack = (2 * 0.85 / 0.15 + 1) * 0.85 * 2.0**-39 * 1.01
    assert sc["l1_error_bound"] >= slack


@pytest.mark.parametrize("weighted,sinks,teleport", [(True, 0, False), (False, 1, False),
                                                     (True, 1, True)])
def test_mixed_closed_form_start(monkeypatch, weighted, sinks, teleport):
    """A mixed solve from a constant start takes its first residual in closed form (d c rowsum(P)
    + b): against the same solve with a first FP64 sweep (NKPR_CLOSED_START=0) it certifies at
    tol, matches the converged oracle, takes the same refinements and one FP64 sweep fewer.  A
    start that is not constant takes the sweep."""
    z, u, v, w = skewed_graph(30000, 250000, 73, weighted=weighted)
    g = oracle.in_lists(z, u, v, w, True)
    t = None
    if teleport:
        rng = np.random.default_rng(9)
        t = np.zeros(z)
        np.add.at(t, rng.choice(z, 40), rng.uniform(0.5, 2.0, 40))
        t /= t.sum()
    ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000,
                                sink_mode=sinks, teleport=t)
    w32 = None if g.in_w is None else g.in_w.astype(np.float32)
    runs = {}
    for flag, start in (("1", "const"), ("0", "const"), ("1", "warm")):
        monkeypatch.setenv("NKPR_CLOSED_START", flag)
        dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), w32, weight64=g.in_w)
        if start == "const":
            x = np.full(z, 1.0 / z)
        else:
            x = np.random.default_rng(10).uniform(0.5, 1.5, z)
            x /= x.sum()
        p = capi.default_params(precision=capi.NKPR_MIXED, tol=1e-10, sink_handling=sinks,
                                stop=capi.NKPR_STOP_L1_ERROR)
        if teleport:
            p.teleport = t.ctypes.data
        st = dg.run(x, p)
        assert st["l1_error_bound"] <= 1e-10, (flag, start)
        assert np.abs(x - ref).sum() <= 1e-10, (flag, start)
        runs[(flag, start)] = st
    closed, swept, warm = runs[("1", "const")], runs[("0", "const")], runs[("1", "warm")]
    assert abs(int(closed["refinements"]) - int(swept["refinements"])) <= 1
    if closed["refinements"] == swept["refinements"]:
        assert closed["sweeps_fp64"] == swept["sweeps_fp64"] - 1
    assert warm["sweeps_fp64"] == warm["refinements"] + 1  # every residual was a sweep


@pytest.mark.parametrize("binned", ["0", "1"])
def test_mixed_solve_timing_stats(monkeypatch, binned):
    """With timing on, a mixed solve reports the device time of its operator applications and
    FP64 sweeps (HIP events; a binned application's events are read when the next one starts or
    when the run ends): both positive, together within the run's wall time, and close per
    application on a second run of the same handle, so that no application's time moves to
    another run; the scores equal those of an untimed run bit for bit."""
    monkeypatch.setenv("NKPR_BINNED", binned)
    z, u, v, w = skewed_graph(60000, 500000, 17, weighted=True)
    g = oracle_graph(z, u, v, w, True)
    dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), g.in_w.astype(np.float32),
                          weight64=g.in_w)
    runs = []
    for timing in (1, 1, 0):
        x = np.full(z, 1.0 / z)
        st = dg.run(x, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-9,
                                           stop=capi.NKPR_STOP_L1_ERROR, timing=timing))
        runs.append((x, st))
    dg.close()
    for x, st in runs[:2]:
        assert st["sweeps_fp32"] > 0 and st["ms_sweeps_fp32"] > 0 and st["ms_sweeps_fp64"] > 0
        assert st["ms_sweeps_fp32"] + st["ms_sweeps_fp64"] <= st["ms_total"]
    per = [st["ms_sweeps_fp32"] / st["sweeps_fp32"] for _, st in runs[:2]]
    assert 0.5 < per[0] / per[1] < 2.0, per
    assert runs[2][1]["ms_sweeps_fp32"] == 0.0
    assert np.array_equal(runs[0][0], runs[2][0]) and np.array_equal(runs[1][0], runs[2][0])


@pytest.mark.parametrize("on_device", [False, True])
def test_constant_start_without_gather_is_bit_identical(on_device):
    """A mixed run from a constant start on a graph whose every slot holds a node checks the
    start once in slot order and never gathers it (its first correction is written as c + e);
    the same graph with a deleted slot appended takes the gathered start.  Both runs give the
    same bits on every node, the deleted slot keeps its value, and the result matches the
    converged oracle."""
    import torch
    z, u, v, w = skewed_graph(20000, 150000, 5, weighted=True)
    ex = np.ones(z + 1, np.uint8)
    ex[z] = 0
    ref, _, _ = oracle.pagerank(oracle.in_lists(z, u, v, w, True), tol=1e-15, norm=oracle.L1,
                                max_iterations=20000)
    out = []
    for zz, exists in ((z, None), (z + 1, ex)):
        g = oracle.in_lists(zz, u, v, w, True, exists=exists)
        dg = capi.DeviceGraph(zz, g.in_ptr, g.in_src.astype(np.int32), g.in_w.astype(np.float32),
                              weight64=g.in_w, exists=exists)
        p = capi.default_params(precision=capi.NKPR_MIXED, tol=1e-10, stop=capi.NKPR_STOP_L1_ERROR,
                                scores_on_device=1 if on_device else 0)
        if on_device:
            x = torch.full((zz,), 1.0 / z, dtype=torch.float64, device="cuda")
            st = dg.run(x, p)
            torch.cuda.synchronize()
            x = x.cpu().numpy()
        else:
            x = np.full(zz, 1.0 / z)
            st = dg.run(x, p)
        dg.close()
        out.append((x, st))
    (a, sa), (b, sb) = out
    assert sa["l1_error_bound"] <= 1e-10
    assert np.abs(a - ref).sum() <= 1e-10
    assert sa["iterations"] == sb["iterations"] and sa["refinements"] == sb["refinements"]
    assert np.array_equal(a, b[:z])
    assert b[z] == 1.0 / z


def test_binned_copies_fall_back_when_memory_is_short(monkeypatch):
    """A device without room for the binned copies (forced from the host: the build's first
    allocation is refused) keeps the pull sweep: the FP64 run still has the oracle's iteration
    count and the mixed run still certifies; later runs on the handle do not retry the build."""
    monkeypatch.setenv("NKPR_BINNED", "1")
    monkeypatch.setenv("NKPR_BINNED64", "1")
    monkeypatch.setenv("NKPR_TEST_BINNED_NOMEM", "1")
    z, u, v, w = skewed_graph(20000, 150000, 68)
    g = oracle.in_lists(z, u, v, w, True)
    dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), g.in_w.astype(np.float32),
                          weight64=g.in_w)
    for _ in range(2):
        x = np.full(z, 1.0 / z)
        st = dg.run(x, capi.default_params())
        ref, it, _ = oracle.pagerank(g)
        assert st["iterations"] == it
        assert np.abs(x - ref).sum() <= 1e-12
    ref2, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000)
    x2 = np.full(z, 1.0 / z)
    st2 = dg.run(x2, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-10,
                                         stop=capi.NKPR_STOP_L1_ERROR))
    assert st2["l1_error_bound"] <= 1e-10
    assert np.abs(x2 - ref2).sum() <= 1e-10
