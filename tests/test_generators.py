"""Workload generators: determinism and the structural properties each one promises."""
import numpy as np

from networkit_amd import generators as gen


def test_rmat_is_deterministic_and_simple():
    a = gen.rmat(12, 8, seed=2)
    b = gen.rmat(12, 8, seed=2)
    c = gen.rmat(12, 8, seed=3)
    assert np.array_equal(a["src"], b["src"]) and np.array_equal(a["colptr"], b["colptr"])
    assert not np.array_equal(a["src"], c["src"])
    z = a["z"]
    rows = np.repeat(np.arange(z), np.diff(a["colptr"].astype(np.int64)))
    assert not np.any(rows == a["src"])                 # no self-loops
    key = rows.astype(np.int64) * z + a["src"]
    assert len(np.unique(key)) == len(key)               # no duplicate arcs
    assert np.all(np.diff(key) > 0)                      # rows ascending, sources ascending
    assert a["weight"].dtype == np.float32
    assert a["weight"].min() >= 0.5 and a["weight"].max() < 1.5
    # weights are exact multiples of 2^-24 (24 random bits)
    assert np.all((a["weight"].astype(np.float64) * 2 ** 24) % 1 == 0)


def test_rmat_chunked_draws_match_single_call():
    L = gen.lib()
    n = 5000
    s1 = np.empty(n, np.int32); d1 = np.empty(n, np.int32)
    L.nkgen_rmat(5, 14, 0, n, .57, .19, .19, s1.ctypes.data, d1.ctypes.data, 0)
    s2 = np.empty(n, np.int32); d2 = np.empty(n, np.int32)
    L.nkgen_rmat(5, 14, 0, 2000, .57, .19, .19, s2.ctypes.data, d2.ctypes.data, 0)
    L.nkgen_rmat(5, 14, 2000, 3000, .57, .19, .19, s2[2000:].ctypes.data, d2[2000:].ctypes.data, 0)
    assert np.array_equal(s1, s2) and np.array_equal(d1, d2)


def test_rmat_quadrant_statistics():
    L = gen.lib()
    n = 200000
    s = np.empty(n, np.int32); d = np.empty(n, np.int32)
    L.nkgen_rmat(9, 1, 0, n, .57, .19, .19, s.ctypes.data, d.ctypes.data, 0)
    top_s = s >> 0 & 1  # least significant bit = last level choice
    top_d = d & 1
    p = [np.mean((top_s == i) & (top_d == j)) for i in (0, 1) for j in (0, 1)]
    assert np.allclose(p, [.57, .19, .19, .05], atol=0.005)


def test_erdos_renyi_density_and_simplicity():
    u, v = gen.erdos_renyi(10000, 1e-3, seed=1)
    pairs = 10000 * 9999 / 2
    assert abs(len(u) - pairs * 1e-3) < 5 * np.sqrt(pairs * 1e-3)
    assert np.all(u < v)
    assert len(np.unique(u * 10000 + v)) == len(u)
    u2, v2 = gen.erdos_renyi(10000, 1e-3, seed=1)
    assert np.array_equal(u, u2) and np.array_equal(v, v2)


def test_barabasi_albert_shape():
    u, v = gen.barabasi_albert(3000, 4, seed=4)
    assert len(u) == (3000 - 4) * 4
    assert np.all(v < u)
    for i in (10, 500, 2999):
        t = v[u == i]
        assert len(t) == 4 and len(np.unique(t)) == 4
    deg = np.bincount(np.concatenate([u, v]))
    assert deg.max() > 10 * np.median(deg)  # heavy tail


def test_csc_from_edges_undirected_matches_networkit_semantics():
    g = gen.csc_from_edges(3, [0, 1, 1], [1, 1, 2], [2.0, 5.0, 7.0], directed=False)
    assert g["colptr"].tolist() == [0, 1, 4, 5]
    assert g["src"].tolist() == [1, 0, 1, 2, 1]
    assert g["weight"].tolist() == [2.0, 2.0, 5.0, 7.0, 7.0]
