"""GPU checks of the teleport vector, nkpr_graph_set_weights and the adjoint (differentiable
PageRank).

Bars:
  teleport, FP64 reference rule: same iteration count as the oracle run with that teleport
      vector and L1 difference <= 1e-12.
  set_weights: the same scores as a graph created with the new weights (L1 <= 1e-12, same
      iteration count).
  adjoint, FP64: gradients equal torch autograd of a dense FP64 solve on the CPU to 1e-9
      relative (2-norm), torch.autograd.gradcheck passes, and central finite differences in FP64
      on R-MAT scale 12 agree to 1e-5 relative per checked edge.
  adjoint, FP32 (restarted FP32 GMRES): within 1e-3 relative (2-norm) of the FP64 gradients.
"""
import numpy as np
import pytest
import torch

import oracle
from graphs import random_graph, skewed_graph
from networkit_amd import autograd as nkad
from networkit_amd import capi
from networkit_amd import generators as gen

pytestmark = pytest.mark.gpu


def device_graph(g):
    w32 = None if g.in_w is None else g.in_w.astype(np.float32)
    return capi.DeviceGraph(g.z, g.in_ptr, g.in_src.astype(np.int32), w32, exists=g.exists,
                            weight64=g.in_w)


def teleport(z, seeds, rng):
    t = np.zeros(z)
    np.add.at(t, seeds, rng.uniform(0.5, 2.0, len(seeds)))
    return t / t.sum()


@pytest.mark.parametrize("sinks", [0, 1])
def test_run_with_teleport_matches_oracle(sinks):
    z, u, v, w = skewed_graph(5000, 40000, 51)
    g = oracle.in_lists(z, u, v, w, True)
    rng = np.random.default_rng(1)
    t = teleport(z, rng.choice(z, 30), rng)
    dg = device_graph(g)
    x = np.full(z, 1.0 / z)
    st = dg.run(x, capi.default_params(sink_handling=sinks, teleport=t.ctypes.data))
    ref, it, _ = oracle.pagerank(g, sink_mode=sinks, teleport=t)
    assert st["iterations"] == it
    assert np.abs(x - ref).sum() <= 1e-12
    # mixed precision with a teleport vector: certified against the converged oracle
    ref2, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=20000,
                                 sink_mode=sinks, teleport=t)
    x2 = np.full(z, 1.0 / z)
    st2 = dg.run(x2, capi.default_params(sink_handling=sinks, teleport=t.ctypes.data,
                                         precision=capi.NKPR_MIXED, tol=1e-9))
    assert st2["l1_error_bound"] <= 1e-9
    assert np.abs(x2 - ref2).sum() <= 1e-9


def test_set_weights_equals_fresh_graph():
    z, u, v, w = skewed_graph(6000, 50000, 52)
    g1 = oracle.in_lists(z, u, v, w, True)
    w2 = np.random.default_rng(2).uniform(0.1, 3.0, len(w))
    g2 = oracle.in_lists(z, u, v, w2, True)
    dg = device_graph(g1)
    dg.set_weights(g2.in_w.astype(np.float32), g2.in_w)
    x = np.full(z, 1.0 / z)
    st = dg.run(x, capi.default_params())
    fresh = device_graph(g2)
    y = np.full(z, 1.0 / z)
    st2 = fresh.run(y, capi.default_params())
    ref, it, _ = oracle.pagerank(g2)
    assert st["iterations"] == st2["iterations"] == it
    assert np.abs(x - y).sum() <= 1e-12 and np.abs(x - ref).sum() <= 1e-12


def dense_grads(z, colptr, src, w, t, c, d=0.85):
    """torch autograd through a dense FP64 solve on the CPU (independent of the adjoint)."""
    wt = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    tt = torch.tensor(t, dtype=torch.float64, requires_grad=True)
    rows = torch.repeat_interleave(torch.arange(z), torch.tensor(np.diff(colptr.astype(np.int64))))
    A = torch.zeros(z, z, dtype=torch.float64).index_put((rows, torch.tensor(src, dtype=torch.int64)), wt, accumulate=True)
    deg = A.sum(0)
    P = A / torch.where(deg > 0, deg, torch.ones_like(deg))
    y = torch.linalg.solve(torch.eye(z, dtype=torch.float64) - d * P, (1 - d) * tt)
    x = y / y.sum()
    L = (torch.tensor(c) * x).sum()
    L.backward()
    return wt.grad.numpy(), tt.grad.numpy(), x.detach().numpy()


def op_for(g, precision=capi.NKPR_FP64, tol=1e-15):
    return nkad.PageRankOperator(g.in_ptr.astype(np.int64), g.in_src.astype(np.int32),
                                 tol=tol, precision=precision)


@pytest.mark.parametrize("precision", [capi.NKPR_FP64, capi.NKPR_MIXED])
def test_adjoint_matches_dense_autograd(precision):
    z, u, v, w = random_graph(300, 1500, 53, directed=True, weighted=True)
    g = oracle.in_lists(z, u, v, w, True)
    rng = np.random.default_rng(3)
    t = teleport(z, rng.choice(z, 20), rng)
    c = rng.normal(size=z)
    gw_ref, gt_ref, x_ref = dense_grads(z, g.in_ptr, g.in_src, g.in_w, t, c)
    op = op_for(g, precision, tol=1e-14 if precision == capi.NKPR_MIXED else 1e-15)
    wt = torch.tensor(g.in_w, dtype=torch.float64, device="cuda", requires_grad=True)
    tt = torch.tensor(t, dtype=torch.float64, device="cuda", requires_grad=True)
    x = op(wt, tt)
    assert np.abs(x.detach().cpu().numpy() - x_ref).sum() <= 1e-12
    L = (x * torch.tensor(c, device="cuda")).sum()
    L.backward()
    gw = wt.grad.cpu().numpy()
    gt = tt.grad.cpu().numpy()
    assert np.linalg.norm(gw - gw_ref) <= 1e-9 * np.linalg.norm(gw_ref)
    assert np.linalg.norm(gt - gt_ref) <= 1e-9 * np.linalg.norm(gt_ref)
    # a sign or index slip in the edge formula would not survive this
    assert np.linalg.norm(gw_ref) > 0 and np.abs(gw_ref).max() > 1e-6


def test_adjoint_gradcheck():
    z, u, v, w = random_graph(14, 45, 54, directed=True, weighted=True, self_loops=False)
    g = oracle.in_lists(z, u, v, w, True)
    op = op_for(g)
    wt = torch.tensor(g.in_w, dtype=torch.float64, device="cuda", requires_grad=True)
    tt = torch.full((z,), 1.0 / z, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: op(a, b), (wt, tt), eps=1e-6, atol=1e-9,
                                    rtol=1e-5)


def test_adjoint_finite_differences_rmat12():
    r = gen.rmat(12, 16, seed=5)
    z = r["z"]
    colptr = r["colptr"].astype(np.int64)
    src = r["src"].astype(np.int32)
    w0 = r["weight"].astype(np.float64)
    op = nkad.PageRankOperator(colptr, src, tol=1e-15, precision=capi.NKPR_FP64)
    rng = np.random.default_rng(5)
    c = torch.tensor(rng.normal(size=z), device="cuda")
    wt = torch.tensor(w0, device="cuda", requires_grad=True)
    L = (op(wt) * c).sum()
    L.backward()
    gw = wt.grad.cpu().numpy()
    idx = rng.choice(len(w0), 40, replace=False)
    for e in idx:
        h = 1e-4 * w0[e]
        wp = w0.copy(); wp[e] += h
        wm = w0.copy(); wm[e] -= h
        lp = float((op(torch.tensor(wp, device="cuda")) * c).sum())
        lm = float((op(torch.tensor(wm, device="cuda")) * c).sum())
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gw[e]) <= 1e-5 * abs(gw[e]) + 1e-11, (e, fd, gw[e])


def test_adjoint_fp32_close_to_fp64():
    r = gen.rmat(14, 16, seed=5)
    colptr = r["colptr"].astype(np.int64)
    src = r["src"].astype(np.int32)
    w = torch.tensor(r["weight"].astype(np.float32), device="cuda")
    c = torch.tensor(np.random.default_rng(6).normal(size=r["z"]), device="cuda")
    grads = {}
    for prec in (capi.NKPR_FP32, capi.NKPR_FP64):
        op = nkad.PageRankOperator(colptr, src, tol=1e-5 if prec == capi.NKPR_FP32 else 1e-14,
                                   precision=prec)
        wt = w.clone().requires_grad_(True)
        L = (op(wt).double() * c).sum()
        L.backward()
        grads[prec] = wt.grad.double().cpu().numpy()
    a, b = grads[capi.NKPR_FP32], grads[capi.NKPR_FP64]
    assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b)


def test_adjoint_host_arrays_and_invalid():
    z, u, v, w = random_graph(400, 2000, 55, directed=True, weighted=True)
    g = oracle.in_lists(z, u, v, w, True)
    dg = device_graph(g)
    x = np.full(z, 1.0 / z)
    st = dg.run(x, capi.default_params(tol=1e-15, norm=capi.NKPR_L1))
    c = np.random.default_rng(7).normal(size=z)
    gw, gt, ast = dg.adjoint(g.in_ptr, g.in_src.astype(np.int32), x, st["score_sum"], c,
                             weight=g.in_w.astype(np.float32), weight64=g.in_w,
                             params=capi.default_params(tol=1e-14))
    gw_ref, gt_ref, _ = dense_grads(z, g.in_ptr, g.in_src, g.in_w, np.full(z, 1.0 / z), c)
    assert np.linalg.norm(gw - gw_ref) <= 1e-9 * np.linalg.norm(gw_ref)
    assert np.linalg.norm(gt - gt_ref) <= 1e-9 * np.linalg.norm(gt_ref)
    with pytest.raises(capi.NkprError):
        dg.adjoint(g.in_ptr, g.in_src.astype(np.int32), x, st["score_sum"], c,
                   weight=g.in_w.astype(np.float32),
                   params=capi.default_params(sink_handling=capi.NKPR_SINK_DISTRIBUTE))
    with pytest.raises(capi.NkprError):
        dg.adjoint(g.in_ptr, g.in_src.astype(np.int32), x, 0.0, c,
                   weight=g.in_w.astype(np.float32))


def test_backward_after_a_later_forward_uses_its_own_weights():
    # op(w1), op(w2), then backward of the first: the handle holds w2 by then, and the gradient
    # of the first call must still be the one of P(w1)
    z, u, v, w = random_graph(250, 1400, 56, directed=True, weighted=True)
    g = oracle.in_lists(z, u, v, w, True)
    rng = np.random.default_rng(8)
    w2 = rng.uniform(0.2, 4.0, len(g.in_w))
    c = rng.normal(size=z)
    uni = np.full(z, 1.0 / z)
    gw1_ref, _, x1_ref = dense_grads(z, g.in_ptr, g.in_src, g.in_w, uni, c)
    gw2_ref, _, x2_ref = dense_grads(z, g.in_ptr, g.in_src, w2, uni, c)
    assert np.linalg.norm(gw1_ref - gw2_ref) > 1e-3 * np.linalg.norm(gw1_ref)
    op = op_for(g)
    ct = torch.tensor(c, device="cuda")
    wt1 = torch.tensor(g.in_w, dtype=torch.float64, device="cuda", requires_grad=True)
    wt2 = torch.tensor(w2, dtype=torch.float64, device="cuda", requires_grad=True)
    x1 = op(wt1)
    x2 = op(wt2)
    assert np.abs(x1.detach().cpu().numpy() - x1_ref).sum() <= 1e-12
    assert np.abs(x2.detach().cpu().numpy() - x2_ref).sum() <= 1e-12
    (x1 * ct).sum().backward()
    gw1 = wt1.grad.cpu().numpy()
    assert np.linalg.norm(gw1 - gw1_ref) <= 1e-9 * np.linalg.norm(gw1_ref)
    (x2 * ct).sum().backward()
    gw2 = wt2.grad.cpu().numpy()
    assert np.linalg.norm(gw2 - gw2_ref) <= 1e-9 * np.linalg.norm(gw2_ref)
