"""Seeded synthetic graphs for the PageRank workloads (wraps libnkgen.so).

Every generator is a pure function of its parameters and seed: the host and the device
produce the same graph.  Graphs are returned as in-lists (CSC of the pull operator, the layout
nkpr_graph_create takes): ``dict(z, colptr, src, weight)`` with numpy arrays on the host or
torch tensors on the GPU.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnkgen.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libnkgen.so is not built: run __graft_entry__.build()")
        from ._runtime import bind_hip_runtime
        bind_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        vp, u64 = ctypes.c_void_p, ctypes.c_uint64
        L.nkgen_rmat.argtypes = [u64, ctypes.c_int, u64, u64, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, vp, vp, ctypes.c_int]
        L.nkgen_rmat.restype = ctypes.c_int
        L.nkgen_edge_weights.argtypes = [u64, vp, vp, u64, vp, ctypes.c_int]
        L.nkgen_edge_weights.restype = ctypes.c_int
        L.nkgen_erdos_renyi.argtypes = [u64, u64, ctypes.c_double, ctypes.c_int, vp, vp,
                                        ctypes.c_int64]
        L.nkgen_erdos_renyi.restype = ctypes.c_int64
        L.nkgen_barabasi_albert.argtypes = [u64, u64, u64, vp, vp, ctypes.c_int64]
        L.nkgen_barabasi_albert.restype = ctypes.c_int64
        L.nkgen_rhg_points.argtypes = [u64, u64, ctypes.c_double, ctypes.c_double, vp, vp, vp]
        L.nkgen_rhg_points.restype = ctypes.c_int
        L.nkgen_rhg.argtypes = [u64, u64, ctypes.c_double, ctypes.c_double, vp, vp,
                                ctypes.c_int64]
        L.nkgen_rhg.restype = ctypes.c_int64
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data if isinstance(a, np.ndarray) else a.data_ptr()


def csc_from_edges(z, u, v, w=None, directed=True):
    """In-lists of the graph NetworKit builds with addEdge(u[i], v[i], w[i]) in order:
    directed -> row v lists u; undirected -> rows u and v list each other, a self-loop once.
    Entries of a row keep insertion order."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    if directed:
        rows, srcs = v, u
        ws = None if w is None else np.asarray(w, dtype=np.float64)
    else:
        # entries in insertion order: edge i adds (row u, src v), then (row v, src u)
        loop = u == v
        rows = np.stack([u, v], axis=1).ravel()
        srcs = np.stack([v, u], axis=1).ravel()
        keep = np.ones(len(rows), bool)
        keep[1::2] = ~loop
        rows, srcs = rows[keep], srcs[keep]
        ws = None
        if w is not None:
            wf = np.asarray(w, np.float64)
            ws = np.stack([wf, wf], axis=1).ravel()[keep]
    order = np.argsort(rows, kind="stable")
    colptr = np.zeros(z + 1, dtype=np.uint64)
    if len(rows):
        colptr[1:] = np.cumsum(np.bincount(rows, minlength=z))
    out = dict(z=int(z), colptr=colptr, src=srcs[order].astype(np.int32),
               weight=None if ws is None else ws[order])
    return out


def rmat(scale, edge_factor=16, seed=2, a=0.57, b=0.19, c=0.19, weighted=True,
         device="cpu"):
    """Directed R-MAT: 2^scale nodes, edge_factor * 2^scale draws, self-loops and duplicate
    arcs removed; weights U[0.5, 1.5) in FP32 (a hash of the arc and the seed).  Rows are
    destinations, sources ascending within a row."""
    z = 1 << scale
    draws = int(edge_factor) << scale
    L = lib()
    if device == "cpu":
        src = np.empty(draws, dtype=np.int32)
        dst = np.empty(draws, dtype=np.int32)
        if L.nkgen_rmat(seed, scale, 0, draws, a, b, c, _ptr(src), _ptr(dst), 0) != 0:
            raise RuntimeError("nkgen_rmat failed")
        key = (dst.astype(np.int64) << 32) | src.astype(np.int64)
        key = key[src != dst]
        key = np.unique(key)
        d = (key >> 32).astype(np.int32)
        s = (key & 0xFFFFFFFF).astype(np.int32)
        colptr = np.zeros(z + 1, dtype=np.uint64)
        colptr[1:] = np.cumsum(np.bincount(d, minlength=z))
        w = None
        if weighted:
            w = np.empty(len(s), dtype=np.float32)
            L.nkgen_edge_weights(seed, _ptr(s), _ptr(d), len(s), _ptr(w), 0)
        return dict(z=z, colptr=colptr, src=s, weight=w)
    import torch
    dev = torch.device(device)
    keys = []
    chunk = 1 << 28
    for first in range(0, draws, chunk):
        cnt = min(chunk, draws - first)
        s = torch.empty(cnt, dtype=torch.int32, device=dev)
        d = torch.empty(cnt, dtype=torch.int32, device=dev)
        torch.cuda.synchronize(dev)
        if L.nkgen_rmat(seed, scale, first, cnt, a, b, c, s.data_ptr(), d.data_ptr(), 1) != 0:
            raise RuntimeError("nkgen_rmat failed on the device")
        k = (d.to(torch.int64) << 32) | s.to(torch.int64)
        keys.append(k[s != d])
        del s, d, k
    key = torch.cat(keys)
    del keys
    key, _ = torch.sort(key)
    keep = torch.ones(key.numel(), dtype=torch.bool, device=dev)
    keep[1:] = key[1:] != key[:-1]
    key = key[keep]
    del keep
    d = (key >> 32).to(torch.int32)
    s = (key & 0xFFFFFFFF).to(torch.int32)
    del key
    colptr = torch.zeros(z + 1, dtype=torch.int64, device=dev)
    colptr[1:] = torch.cumsum(torch.bincount(d, minlength=z), 0)
    w = None
    if weighted:
        w = torch.empty(s.numel(), dtype=torch.float32, device=dev)
        torch.cuda.synchronize(dev)
        if L.nkgen_edge_weights(seed, s.data_ptr(), d.data_ptr(), s.numel(), w.data_ptr(), 1):
            raise RuntimeError("nkgen_edge_weights failed on the device")
    del d
    return dict(z=z, colptr=colptr, src=s, weight=w)


def erdos_renyi(n, p, seed=1, directed=False):
    """G(n, p) edge list (u, v) in generation order."""
    L = lib()
    m = L.nkgen_erdos_renyi(seed, n, p, int(directed), None, None, 0)
    u = np.empty(max(m, 1), dtype=np.int64)
    v = np.empty(max(m, 1), dtype=np.int64)
    got = L.nkgen_erdos_renyi(seed, n, p, int(directed), _ptr(u), _ptr(v), m)
    assert got == m
    return u[:m], v[:m]


def barabasi_albert(n, k, seed=4):
    """Barabasi-Albert edge list (u, v): node i >= k attaches k distinct older targets."""
    L = lib()
    m = L.nkgen_barabasi_albert(seed, n, k, None, None, 0)
    u = np.empty(max(m, 1), dtype=np.int64)
    v = np.empty(max(m, 1), dtype=np.int64)
    got = L.nkgen_barabasi_albert(seed, n, k, _ptr(u), _ptr(v), m)
    assert got == m
    return u[:m], v[:m]


def rhg_points(n, avg_degree=20.0, gamma=3.0, seed=3):
    """Host copy of the threshold random hyperbolic graph's points: (radius r[n], angle
    theta[n], disk radius R)."""
    r = np.empty(n, dtype=np.float64)
    t = np.empty(n, dtype=np.float64)
    R = ctypes.c_double(0.0)
    if lib().nkgen_rhg_points(seed, n, avg_degree, gamma, _ptr(r), _ptr(t), ctypes.byref(R)):
        raise ValueError("bad random hyperbolic graph parameters")
    return r, t, R.value


def rhg(n, avg_degree=20.0, gamma=3.0, seed=3):
    """Undirected threshold random hyperbolic graph (temperature 0) built on the GPU: n points,
    power-law exponent gamma, average degree ~avg_degree (asymptotic disk radius).  Returns the
    in-lists as torch CUDA tensors: dict(z, colptr int64[n+1], src int32[m], weight None)."""
    import torch
    m = lib().nkgen_rhg(seed, n, avg_degree, gamma, None, None, 0)
    if m < 0:
        raise RuntimeError("nkgen_rhg failed (%d)" % m)
    colptr = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    src = torch.empty(max(m, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    got = lib().nkgen_rhg(seed, n, avg_degree, gamma, _ptr(colptr), _ptr(src), m)
    if got != m:
        raise RuntimeError("nkgen_rhg failed (%d)" % got)
    return dict(z=n, colptr=colptr, src=src[:m], weight=None)
