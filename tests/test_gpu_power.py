"""GPU parity of EigenvectorCentrality and KatzCentrality (nkpr_power_run) against the oracle
and the reference's known answers.

Bars: the oracle's iteration count and an L1 difference <= 1e-12 relative to the L1 norm of
the scores (summation order only: both loops are FP64); the 8-node eigenvector values to 1e-4
(CentralityGTest.cpp:541-578) and wiki-Vote Katz top-1 = node 699 with alpha 5e-4 (:216-225);
deleted slots hold 1 as in the reference.
"""
import numpy as np
import pytest

import networkit_amd as nk
import oracle
from graphs import eight_node, random_graph, skewed_graph, wiki_vote
from networkit_amd import capi
from networkit_amd import generators as gen

pytestmark = pytest.mark.gpu

EIGHT_EV = [0.2254, 0.1503, 0.5290, 0.4508, 0.3006, 0.5290, 0.2254, 0.1503]
CAP = 100000  # the oracle's loop cap (the reference has none): a case that does not converge fails


def build(z, u, v, w, directed, remove=()):
    G = nk.Graph(z, w is not None, directed)
    G.addEdges(u, v, w)
    for x in remove:
        G.removeNode(x)
    return G


def oracle_lists(z, u, v, w, directed, remove=(), outgoing=False):
    keep = np.ones(len(u), bool)
    for x in remove:
        keep &= (u != x) & (v != x)
    exists = None
    if remove:
        exists = np.ones(z, np.uint8)
        exists[list(remove)] = 0
    a, b = (v, u) if (outgoing and directed) else (u, v)
    return oracle.in_lists(z, a[keep], b[keep], None if w is None else w[keep], directed,
                           exists=exists)


def close(got, ref):
    return np.abs(got - ref).sum() <= 1e-12 * max(1.0, np.abs(ref).sum())


def check_eigen(z, u, v, w, directed, remove=(), tol=1e-9):
    G = build(z, u, v, w, directed, remove)
    ev = nk.centrality.EigenvectorCentrality(G, tol)
    ev.run()
    got = np.array(ev.scores())
    ref, it = oracle.eigenvector(oracle_lists(z, u, v, w, directed, remove), tol,
                                 max_iterations=CAP)
    assert it < CAP
    assert ev.stats()["iterations"] == it
    assert close(got, ref), np.abs(got - ref).sum()
    return ev, got


def check_katz(z, u, v, w, directed, remove=(), alpha=0, beta=0.1, tol=1e-8, outgoing=False):
    G = build(z, u, v, w, directed, remove)
    kc = nk.centrality.KatzCentrality(G, alpha, beta, tol)
    if outgoing:
        kc.edgeDirection = nk.centrality.EdgeDirection.outEdges
    kc.run()
    got = np.array(kc.scores())
    og = oracle_lists(z, u, v, w, directed, remove, outgoing)
    a = alpha if alpha else oracle.katz_default_alpha(oracle_lists(z, u, v, w, directed, remove))
    ref, it = oracle.katz(og, alpha=a, beta=beta, tol=tol, max_iterations=CAP)
    assert it < CAP
    assert kc.stats()["iterations"] == it
    assert close(got, ref), np.abs(got - ref).sum()
    return kc, got


def test_eigenvector_eight_node_known_answers():
    z, u, v, w = eight_node()
    ev, got = check_eigen(z, u, v, w, False, tol=1e-8)
    assert np.allclose(got, EIGHT_EV, atol=1e-4)
    # the reference's C++ default tolerance
    G = build(z, u, v, w, False)
    c = nk.centrality.EigenvectorCentrality(G, 1e-8)
    c.run()
    assert np.allclose(c.scores(), EIGHT_EV, atol=1e-4)


def test_katz_wiki_vote_known_answer():
    z, u, v = wiki_vote(False)  # SNAPGraphReader{} reads undirected
    kc, got = check_katz(z, u, v, None, False, alpha=5e-4)
    assert kc.ranking()[0][0] == 699


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("weighted", [True, False])
def test_eigenvector_random_graphs(directed, weighted):
    z, u, v, w = random_graph(4000, 40000, 71, directed=directed, weighted=weighted)
    check_eigen(z, u, v, w, directed)


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("weighted", [True, False])
def test_katz_random_graphs(directed, weighted):
    z, u, v, w = random_graph(4000, 40000, 72, directed=directed, weighted=weighted)
    check_katz(z, u, v, w, directed)
    # a tighter tolerance, still well above the rounding noise of the norm (||x|| is about 90
    # here: the norm itself is only known to about 1e-12, so tol = 1e-12 would let the summation
    # order decide the last iteration)
    check_katz(z, u, v, w, directed, alpha=0.6 / 20, beta=1.0, tol=1e-10)


def test_katz_out_edges_direction():
    # unweighted, so that the default alpha 1 / (1 + max out-degree) stays below
    # 1 / spectral radius in both directions
    z, u, v, _ = skewed_graph(3000, 30000, 73, weighted=False)
    kin, a = check_katz(z, u, v, None, True)
    kout, b = check_katz(z, u, v, None, True, outgoing=True)
    assert np.abs(a - b).sum() > 1e-3  # the two directions differ on a directed graph


def test_deleted_nodes_keep_one():
    z, u, v, w = random_graph(2000, 12000, 74, directed=True, weighted=True)
    dead = (3, 77, 1999)
    _, got = check_eigen(z, u, v, w, True, remove=dead)
    assert all(got[d] == 1.0 for d in dead)
    _, got = check_katz(z, u, v, w, True, remove=dead)
    assert all(got[d] == 1.0 for d in dead)


@pytest.fixture
def sliced(monkeypatch):
    monkeypatch.setenv("NKPR_SLICES", "8")
    monkeypatch.setenv("NKPR_DIRECT_MAX_DEG", "2")


def test_sliced_layout_and_split_rows(sliced):
    # heavy rows (split chunks) and the source-sliced schedule
    rng = np.random.default_rng(75)
    z = 60000
    u = np.concatenate([np.arange(1, 30001), rng.integers(0, z, 300000)])
    v = np.concatenate([np.zeros(30000, np.int64), rng.integers(0, z, 300000)])
    w = rng.uniform(0.5, 1.5, len(u))
    check_eigen(z, u, v, w, True)
    check_katz(z, u, v, None, False)


def test_rmat_eigenvector_and_katz():
    r = gen.rmat(16, 16, seed=6)
    z = r["z"]
    rows = np.repeat(np.arange(z), np.diff(r["colptr"].astype(np.int64)))
    w = r["weight"].astype(np.float64)
    check_eigen(z, r["src"], rows, w, True)
    check_katz(z, r["src"], rows, w, True)


def test_capi_power_rejects_and_diverges():
    import torch
    z, u, v, w = random_graph(500, 3000, 76, directed=True, weighted=True)
    og = oracle.in_lists(z, u, v, w, True)
    # a handle created with the real out-degrees folds w / deg: rejected
    dg = capi.DeviceGraph(z, og.in_ptr, og.in_src.astype(np.int32), og.in_w.astype(np.float32),
                          weight64=og.in_w)
    with pytest.raises(capi.NkprError) as e:
        dg.power_run(np.zeros(z))
    assert e.value.code == capi.NKPR_ERR_INVALID and "out_wdeg" in str(e.value)
    ones = np.ones(z)
    dg1 = capi.DeviceGraph(z, og.in_ptr, og.in_src.astype(np.int32), og.in_w.astype(np.float32),
                           weight64=og.in_w, out_wdeg=ones)
    # device scores give the host result
    x = np.zeros(z)
    st = dg1.power_run(x, kind=capi.NKPR_KATZ, alpha=0.01, beta=0.1)
    xd = torch.zeros(z, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    dg1.power_run(xd, kind=capi.NKPR_KATZ, alpha=0.01, beta=0.1, scores_on_device=1)
    assert np.array_equal(xd.cpu().numpy(), x)
    ref, it = oracle.katz(og, alpha=0.01, beta=0.1, max_iterations=CAP)
    assert st["iterations"] == it and close(x, ref)
    # alpha far above 1 / spectral radius: an error instead of an endless loop
    with pytest.raises(capi.NkprError) as e:
        dg1.power_run(x, kind=capi.NKPR_KATZ, alpha=10.0)
    assert "diverged" in str(e.value)
    with pytest.raises(capi.NkprError):
        dg1.power_run(x, kind=capi.NKPR_KATZ, alpha=-1.0)
    # the handle still works
    st = dg1.power_run(x, kind=capi.NKPR_EIGENVECTOR, tol=1e-10)
    ref, it = oracle.eigenvector(og, 1e-10, max_iterations=CAP)
    assert st["iterations"] == it and close(x, ref)
