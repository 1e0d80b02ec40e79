"""Closed-form and brute-force PageRank references used to pin the oracle and the GPU path.

These do not restate the reference's loop: they solve the fixed point it converges to.

Non-normalised PageRank (networkit/cpp/centrality/PageRank.cpp:56-80) is the solution of
    (I - d P) x = (1 - d)/n 1,        P[u, v] = w_vu / deg_v  (v -> u an edge, deg_v > 0)
rescaled to sum 1.  Mass sent to a node without out-edges is lost; with sink distribution the
operator gains the rank-1 term (d/n) 1 s^T (s = sink indicator), which this module folds in with
the Sherman-Morrison formula.  ``normalized=True`` divides by (1/n)((1 - d) + d * sum_sinks x)
instead (PageRank.cpp:81-92).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _operator(g):
    """Return (P as CSR over existing nodes, map slot->compact, sink mask, n)."""
    z = g.z
    exists = np.ones(z, bool) if g.exists is None else g.exists.astype(bool)
    idx = -np.ones(z, np.int64)
    idx[exists] = np.arange(int(exists.sum()))
    n = int(exists.sum())
    ptr = g.in_ptr.astype(np.int64)
    src = g.in_src.astype(np.int64)
    w = np.ones(len(src)) if g.in_w is None else g.in_w.astype(np.float64)
    rows = np.repeat(np.arange(z), np.diff(ptr))
    deg = np.bincount(src, weights=w, minlength=z)
    cnt = np.bincount(src, minlength=z)
    vals = w / deg[src]
    P = sp.csr_matrix((vals, (idx[rows], idx[src])), shape=(n, n))
    sinks = (cnt == 0)[exists]
    return P, idx, exists, sinks, n


def solve_fixed_point(g, damp=0.85, sink_mode=0, teleport=None):
    """Unnormalised fixed point x over existing nodes (compact order).  ``teleport`` (length
    n, compact order) replaces the uniform (1-d)/n vector when given."""
    P, idx, exists, sinks, n = _operator(g)
    A = (sp.identity(n, format="csc") - damp * P).tocsc()
    b = np.full(n, (1.0 - damp) / n) if teleport is None else np.asarray(teleport, float)
    lu = spla.splu(A)
    x = lu.solve(b)
    if sink_mode == 1 and sinks.any():
        c = np.full(n, damp / n)
        y = lu.solve(c)
        x = x + y * (sinks @ x) / (1.0 - sinks @ y)
    return x, exists, sinks, n


def pagerank_closed_form(g, damp=0.85, sink_mode=0, normalized=False):
    """Final reference scores (length z; deleted slots hold 0 here)."""
    x, exists, sinks, n = solve_fixed_point(g, damp, sink_mode)
    if normalized:
        x = x / ((1.0 / n) * ((1.0 - damp) + damp * x[sinks].sum()))
    else:
        x = x / x.sum()
    out = np.zeros(g.z)
    out[exists] = x
    return out


def dense_power_iteration(g, damp=0.85, sink_mode=0, iters=2000):
    """Brute-force FP64 power iteration on a dense matrix (tiny graphs only)."""
    P, idx, exists, sinks, n = _operator(g)
    M = damp * P.toarray()
    if sink_mode == 1:
        M = M + np.outer(np.full(n, damp / n), sinks.astype(float))
    x = np.full(n, 1.0 / n)
    for _ in range(iters):
        x = M @ x + (1.0 - damp) / n
    out = np.zeros(g.z)
    out[exists] = x / x.sum()
    return out
