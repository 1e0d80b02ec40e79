"""PageRank throughput on one MI355X: mixed precision (FP32 sweeps + FP64 refinement) against
plain FP64 at equal certified L1 error, on a directed R-MAT graph.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 26] [--no-cpu]

One step = one complete PageRank solve (damp 0.85, certified L1 error <= 1e-8 on the
sum-normalised scores) of the graph already resident in HBM, from the uniform start vector.
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes(m, n, weighted, kind):
    """Bytes one sweep of the reference loop (PageRank.cpp:56-66 over a CSC) moves, each array
    element once and one gather per in-edge, as the benchmark's definition counts them:
    fp32 plain:  per edge src 4 + w 4 + gather 4;  per node colptr 8 + x 4 + x' 4 + 1/deg 4
                 + prescaled x 4  (12 m + 24 n)
    fp64 plain:  per edge src 4 + w 8 + gather 8;  per node colptr 8 + x 8 + x' 8 + 1/deg 8
                 + prescaled x 8  (20 m + 40 n)
    unweighted:  no w stream (fp32 8 m + 24 n, fp64 12 m + 40 n)
    fp32corr:    fp32 plain + the right-hand side 4 per node
    fp64resid:   fp64 plain + the FP32 residual write 4 per node
    The folded layout of this library (w/deg stored per edge) skips the two per-node streams of
    1/deg and the prescaled x on weighted graphs, so on them it moves fewer node bytes than
    counted here (12 m + 16 n in FP32)."""
    e = {"fp32": 12, "fp32corr": 12, "fp64": 20}.get(kind, 20)
    if not weighted:
        e -= 4 if kind.startswith("fp32") else 8
    v = {"fp32": 24, "fp32corr": 28, "fp64": 40}.get(kind, 44)
    return e * m + v * n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--sweeps-only", type=int, default=0,
                    help="time N plain FP32 and N plain FP64 sweeps only (profiling runs)")
    ap.add_argument("--workload", choices=["pagerank", "rhg", "ppr", "graphs", "adjoint",
                                           "algebraic", "apr", "nkbg", "power"],
                    default="pagerank",
                    help="pagerank: one solve on R-MAT (the benchmark line); rhg: the same on "
                         "an undirected random hyperbolic graph; ppr: batched personalised "
                         "PageRank on a Barabasi-Albert graph; graphs: one batched run over 512 "
                         "Erdos-Renyi graphs of uneven size; adjoint: forward and backward "
                         "of differentiable PageRank on R-MAT scale 20; algebraic: "
                         "AlgebraicPageRank on NetworKit's CSR matrix of the R-MAT graph; apr: "
                         "ApproximatePageRank pushes on the random hyperbolic graph; nkbg: a "
                         "NetworKit binary file of the R-MAT graph read into device lists; "
                         "power: Eigenvector and Katz centrality on the R-MAT graph")
    ap.add_argument("--rhg-nodes", type=int, default=10_000_000)
    ap.add_argument("--apr-eps", type=float, default=1e-9)
    ap.add_argument("--nkbg-scale", type=int, default=24)
    ap.add_argument("--nkbg-dir", default="/tmp")
    ap.add_argument("--ppr-nodes", type=int, default=2_000_000)
    ap.add_argument("--ppr-columns", type=int, default=256)
    ap.add_argument("--graphs", type=int, default=512)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ppr":
        bench_ppr(args)
        return
    if args.workload == "graphs":
        bench_graphs(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "algebraic":
        bench_algebraic(args)
        return
    if args.workload == "apr":
        bench_apr(args)
        return
    if args.workload == "nkbg":
        bench_nkbg(args)
        return
    if args.workload == "power":
        bench_power(args)
        return

    import torch

    from networkit_amd import capi
    from networkit_amd import generators as gen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.time()
    if args.workload == "rhg":
        g = gen.rhg(args.rhg_nodes, 20.0, 3.0, seed=3)
        workload = ("PageRank solve to certified L1 error <= %g (damp 0.85, no sink handling) of "
                    "an undirected threshold random hyperbolic graph, n=%d, average degree 20, "
                    "power-law exponent 3, temperature 0, unit weights, seed 3"
                    % (args.tol, args.rhg_nodes))
    else:
        g = gen.rmat(args.scale, args.edge_factor, seed=2, device="cuda")
        workload = ("PageRank solve to certified L1 error <= %g (damp 0.85, no sink handling) of "
                    "a directed R-MAT graph, scale %d, edge factor %d, a/b/c/d=.57/.19/.19/.05, "
                    "duplicates and self-loops removed, FP32 weights U(0.5,1.5), seed 2"
                    % (args.tol, args.scale, args.edge_factor))
    weighted = g["weight"] is not None
    torch.cuda.synchronize()
    z = g["z"]
    m = g["src"].numel()
    log("%s graph: z=%d m=%d generated in %.1f s" % (args.workload, z, m, time.time() - t0))
    t0 = time.time()
    dg = capi.DeviceGraph(z, g["colptr"], g["src"], g["weight"], device=True)
    info = dg.info()
    n = info["n"]
    log("device graph: n=%d m=%d sinks=%d max_in_degree=%d items=%d prep %.0f ms (wall %.1f s)"
        % (n, info["m"], info["sinks"], info["max_in_degree"], info["work_items"],
           info["prep_ms"], time.time() - t0))
    stream = torch.cuda.ExternalStream(dg.stream())
    x = torch.empty(z, dtype=torch.float64, device=dev)

    def solve(precision, tol, timing=1):
        # torch's own stream may still read x (clones, error sums): order it before the refill
        torch.cuda.current_stream().synchronize()
        with torch.cuda.stream(stream):
            x.fill_(1.0 / z)
        p = capi.default_params(scores_on_device=1, precision=precision, tol=tol,
                                stop=capi.NKPR_STOP_L1_ERROR, timing=timing)
        return dg.run(x, p)

    if args.sweeps_only:
        # reference copy kernel of known size (1 GiB read + 1 GiB written) for scaling the
        # FETCH_SIZE / WRITE_SIZE counters of a profiled run
        a_src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        a_dst = torch.empty_like(a_src)
        a_dst.copy_(a_src)
        torch.cuda.synchronize()
        # the streaming rate a plain device copy reaches here, for comparison with the passes
        t0e, t1e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0e.record()
        for _ in range(4):
            a_dst.copy_(a_src)
        t1e.record()
        torch.cuda.synchronize()
        cms = t0e.elapsed_time(t1e) / 4
        log("copy 1 GiB: %.3f ms = %.0f GB/s (read + write)" % (cms, 2 * 2 ** 30 / (cms / 1e3) / 1e9))
        del a_src, a_dst
        # plain FP32 and FP64 pull sweeps, then the FP64 sweeps as the library runs them on
        # this graph (binned from 2^26 edges on)
        env64 = os.environ.get("NKPR_BINNED64")
        for name, prec, per_edge, per_node, pull in (("fp32 pull", capi.NKPR_FP32, 12, 16, True),
                                                      ("fp64 pull", capi.NKPR_FP64, 20, 24, True),
                                                      ("fp64", capi.NKPR_FP64, 20, 24, False)):
            if pull:
                os.environ["NKPR_BINNED64"] = "0"
            elif env64 is None:
                os.environ.pop("NKPR_BINNED64", None)
            else:
                os.environ["NKPR_BINNED64"] = env64
            torch.cuda.current_stream().synchronize()
            with torch.cuda.stream(stream):
                x.fill_(1.0 / z)
            p = capi.default_params(scores_on_device=1, precision=prec, tol=0.0,
                                    max_iterations=args.sweeps_only, timing=1)
            st = dg.run(x, p)
            k = st["sweeps_fp32"] + st["sweeps_fp64"]
            ms = (st["ms_sweeps_fp32"] + st["ms_sweeps_fp64"]) / k
            gbs = (per_edge * m + per_node * n) / (ms / 1e3) / 1e9
            log("%s: %d sweeps, %.3f ms/sweep, %.0f GB/s algorithmic, %.3g edge-updates/s" % (
                name, k, ms, gbs, m / (ms / 1e3)))
        # one mixed solve: its FP32 sweeps are the operator applications of the GMRES corrections
        st = solve(capi.NKPR_MIXED, args.tol)
        ms = st["ms_sweeps_fp32"] / st["sweeps_fp32"]
        log("operator (%s): %d applications, %.3f ms each, %.0f GB/s algorithmic" % (
            operator_name(m), st["sweeps_fp32"], ms,
            algorithmic_bytes(m, n, weighted, "fp32") / (ms / 1e3) / 1e9))
        return

    # the reference's own stopping rule (L2 step <= 1e-8, FP64): its sweep count sizes the
    # CPU baseline's estimate of run() (not timed)
    torch.cuda.current_stream().synchronize()
    with torch.cuda.stream(stream):
        x.fill_(1.0 / z)
    its_rule = dg.run(x, capi.default_params(scores_on_device=1, precision=capi.NKPR_FP64,
                                             tol=1e-8))["iterations"]
    # tightly converged FP64 reference (not timed)
    st = solve(capi.NKPR_FP64, 1e-13)
    ref = x.clone()
    torch.cuda.synchronize()
    log("reference: %d FP64 sweeps, certified L1 bound %.2e (reference rule: %d sweeps)" % (
        st["iterations"], st["l1_error_bound"], its_rule))

    def timed(precision, steps, warmup):
        for _ in range(warmup):
            solve(precision, args.tol)
        torch.cuda.synchronize()
        stats = []
        t = time.perf_counter()
        for _ in range(steps):
            stats.append(solve(precision, args.tol))
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, stats

    sec_mixed, st_mixed = timed(capi.NKPR_MIXED, args.steps, args.warmup)
    err_mixed = float((x - ref).abs().sum())
    x_mixed = x.clone()
    torch.cuda.synchronize()
    sec_fp64, st_fp64 = timed(capi.NKPR_FP64, args.steps, args.warmup)
    err_fp64 = float((x - ref).abs().sum())
    # top-1000 ranking agreement of the mixed result with the reference
    top_ref = torch.argsort(ref, descending=True, stable=True)[:1000]
    top_mix = torch.argsort(x_mixed, descending=True, stable=True)[:1000]
    mism = (top_ref != top_mix).nonzero().flatten()
    tie_ok = bool(all(abs(float(ref[top_ref[i]] - ref[top_mix[i]])) <= 1e-12 for i in mism))

    s0 = st_mixed[-1]
    sweeps_mixed = s0["sweeps_fp32"] + s0["sweeps_fp64"]
    sweeps_fp64 = st_fp64[-1]["sweeps_fp64"]
    value = m * sweeps_mixed / sec_mixed
    # roofline of the dominant kernel: the FP32 sweep, timed with HIP events on its stream
    # every FP32 sweep of the mixed path is an operator application y = d P v (GMRES)
    n_f32 = sum(s["sweeps_fp32"] for s in st_mixed)
    ms_f32 = sum(s["ms_sweeps_fp32"] for s in st_mixed)
    bytes_f32 = n_f32 * algorithmic_bytes(m, n, weighted, "fp32")
    achieved = bytes_f32 / (ms_f32 / 1e3) / 1e9
    ms64 = sum(s["ms_sweeps_fp64"] for s in st_fp64)
    n64 = sum(s["sweeps_fp64"] for s in st_fp64)
    achieved64 = n64 * algorithmic_bytes(m, n, weighted, "fp64") / (ms64 / 1e3) / 1e9
    log("mixed: %.1f ms/solve, %d fp32 (GMRES) + %d fp64 sweeps, %d refinements, L1 err %.2e; "
        "fp32 sweep %.3f ms = %.0f GB/s" % (sec_mixed * 1e3, s0["sweeps_fp32"],
                                            s0["sweeps_fp64"], s0["refinements"],
                                            err_mixed, ms_f32 / n_f32, achieved))
    log("fp64: %.1f ms/solve, %d sweeps, L1 err %.2e; fp64 sweep %.3f ms = %.0f GB/s" % (
        sec_fp64 * 1e3, sweeps_fp64, err_fp64, ms64 / n64, achieved64))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_run_baseline(host_in_lists(g["colptr"], g["src"], g["weight"]), m,
                               args.cpu_seconds, its_rule, "the same graph")
    traffic, traffic_src = measured_traffic(m, z, args.workload)
    out = {
        "metric": "PageRank edge-updates/s & %HBM roofline, mixed vs FP64 at equal L1 error",
        "value": value,
        "unit": "edge-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "n": z, "m": m, "sinks": info["sinks"], "parallelism": "single GPU",
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "FP32 operator application y = d P v (GMRES): " + operator_name(m),
                     "bytes_per_sweep": algorithmic_bytes(m, n, weighted, "fp32"),
                     "ms_per_sweep": ms_f32 / n_f32},
        "cpu_baseline": cpu,
        "mixed": {"ms_per_solve": sec_mixed * 1e3, "sweeps_fp32": s0["sweeps_fp32"],
                  "refinements": s0["refinements"],
                  "sweeps_fp64": s0["sweeps_fp64"], "l1_error": err_mixed,
                  "l1_error_bound": s0["l1_error_bound"], "top1000_same": tie_ok},
        "fp64": {"ms_per_solve": sec_fp64 * 1e3, "sweeps": sweeps_fp64, "l1_error": err_fp64,
                 "edge_updates_per_s": m * sweeps_fp64 / sec_fp64,
                 "ms_per_sweep": ms64 / n64, "achieved_GBs": achieved64},
        "speedup_vs_fp64": sec_fp64 / sec_mixed,
        "prep_ms": info["prep_ms"],
    }
    print(json.dumps(out), flush=True)


def spmm_bytes(m, n, k, weighted, dbl):
    """Algorithmic bytes of one batched sweep over k columns: per in-edge the source id, its
    coefficient (w/deg, or 1/deg of the source when unweighted) and one gathered row of k
    values; per node the row offset, the old and the new row of k values."""
    e = 8 if dbl else 4
    return m * (4 + e + e * k) + n * (4 + 2 * e * k)


def bench_ppr(args):
    """Batched personalised PageRank: Barabasi-Albert graph (n nodes, 8 attachments per node,
    undirected, unweighted), k teleport distributions each uniform over a random seed set of
    U{1..1000} nodes, seed 4; mixed precision to a certified L1 error <= tol per column next
    to the plain FP64 batch at the same certified error."""
    import torch

    from networkit_amd import capi
    from networkit_amd import generators as gen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    nn, k = args.ppr_nodes, args.ppr_columns
    t0 = time.time()
    u, v = gen.barabasi_albert(nn, 8, seed=4)
    # undirected in-lists (NetworKit adjacency: edge i adds v to row u, then u to row v; BA has
    # no self-loops), sorted stably by row on the GPU
    tu = torch.from_numpy(u).to(dev)
    tv = torch.from_numpy(v).to(dev)
    rows = torch.stack([tu, tv], dim=1).reshape(-1)
    srcs = torch.stack([tv, tu], dim=1).reshape(-1)
    order = torch.argsort(rows, stable=True)
    colptr = torch.zeros(nn + 1, dtype=torch.int64, device=dev)
    colptr[1:] = torch.cumsum(torch.bincount(rows, minlength=nn), 0)
    c = {"colptr": colptr, "src": srcs[order].to(torch.int32).contiguous()}
    del tu, tv, rows, srcs, order
    m = int(colptr[-1])
    rng = np.random.default_rng(4)
    seeds = [rng.choice(nn, int(rng.integers(1, 1001)), replace=False) for _ in range(k)]
    log("BA n=%d m=%d arcs, %d seed sets, built in %.1f s" % (nn, m, k, time.time() - t0))
    torch.cuda.synchronize()
    dg = capi.DeviceGraph(nn, c["colptr"], c["src"], device=True)
    stream = torch.cuda.ExternalStream(dg.stream())
    x = torch.empty((k, nn), dtype=torch.float64, device=dev)

    def solve(precision, tol):
        torch.cuda.current_stream().synchronize()
        with torch.cuda.stream(stream):
            x.fill_(1.0 / nn)
        p = capi.default_params(scores_on_device=1, precision=precision, tol=tol,
                                stop=capi.NKPR_STOP_L1_ERROR, timing=1)
        _, st, its, bnd = dg.ppr_batch(seeds, scores=x, params=p)
        return st, bnd

    # per-column sweep counts of the reference rule (FP64, L2 step <= 1e-8; not timed)
    torch.cuda.current_stream().synchronize()
    with torch.cuda.stream(stream):
        x.fill_(1.0 / nn)
    _, _, its_rule, _ = dg.ppr_batch(seeds, scores=x, params=capi.default_params(
        scores_on_device=1, precision=capi.NKPR_FP64, tol=1e-8))
    solve(capi.NKPR_FP64, 1e-13)
    ref = x.clone()
    torch.cuda.synchronize()

    def timed(precision):
        for _ in range(args.warmup):
            solve(precision, args.tol)
        torch.cuda.synchronize()
        t = time.perf_counter()
        sts = [solve(precision, args.tol) for _ in range(args.steps)]
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t) / args.steps
        err = float((x - ref).abs().sum(dim=1).max())
        return sec, sts, err

    cpu = None
    if not args.no_cpu:
        cpu = cpu_ppr_baseline(host_in_lists(c["colptr"], c["src"]), nn, m, seeds, its_rule,
                               args.cpu_seconds)
    sec_m, st_m, err_m = timed(capi.NKPR_MIXED)
    sec_d, st_d, err_d = timed(capi.NKPR_FP64)
    s0 = st_m[0][0]
    d0 = st_d[0][0]
    b32 = spmm_bytes(m, nn, k, False, False)
    b64 = spmm_bytes(m, nn, k, False, True)
    ms32 = sum(st["ms_sweeps_fp32"] for st, _ in st_m) / sum(st["sweeps_fp32"] for st, _ in st_m)
    ms64 = sum(st["ms_sweeps_fp64"] for st, _ in st_d) / sum(st["sweeps_fp64"] for st, _ in st_d)
    ach32 = b32 / (ms32 / 1e3) / 1e9
    ach64 = b64 / (ms64 / 1e3) / 1e9
    log("mixed: %.1f ms, %d fp32 + %d fp64 sweeps, max L1 err %.2e; fp32 SpMM %.3f ms = %.0f GB/s "
        "(tile %d, heavy %d)" % (sec_m * 1e3, s0["sweeps_fp32"], s0["sweeps_fp64"], err_m, ms32,
                                 ach32, s0["tile_columns"], s0["heavy_degree"]))
    log("fp64: %.1f ms, %d sweeps, max L1 err %.2e; fp64 SpMM %.3f ms = %.0f GB/s" % (
        sec_d * 1e3, d0["sweeps_fp64"], err_d, ms64, ach64))
    sweeps_m = s0["sweeps_fp32"] + s0["sweeps_fp64"]
    out = {
        "metric": "batched personalised PageRank column-edge-updates/s, mixed vs FP64 at equal "
                  "L1 error",
        "value": m * k * sweeps_m / sec_m, "unit": "column-edge-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec_m * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": "batched personalised PageRank, %d teleport distributions "
                               "(uniform over U{1..1000} random seeds), certified L1 error <= "
                               "%g per column, Barabasi-Albert n=%d, 8 attachments, undirected, "
                               "seed 4" % (k, args.tol, nn), "n": nn, "m": m, "columns": k},
        "roofline": {"bound": "hbm", "achieved": ach32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach32 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "FP32 SpMM sweep", "bytes_per_sweep": b32,
                     "ms_per_sweep": ms32},
        "cpu_baseline": cpu,
        "mixed": {"ms_per_solve": sec_m * 1e3, "sweeps_fp32": s0["sweeps_fp32"],
                  "sweeps_fp64": s0["sweeps_fp64"], "refinements": s0["refinements"],
                  "max_l1_error": err_m, "max_error_bound": s0["max_error_bound"],
                  "tile_columns": s0["tile_columns"], "heavy_degree": s0["heavy_degree"]},
        "fp64": {"ms_per_solve": sec_d * 1e3, "sweeps": d0["sweeps_fp64"], "max_l1_error": err_d,
                 "ms_per_sweep": ms64, "achieved_GBs": ach64,
                 "tile_columns": d0["tile_columns"]},
        "speedup_vs_fp64": sec_d / sec_m,
    }
    print(json.dumps(out), flush=True)


def bench_graphs(args):
    """A batch of graphs in one handle: G undirected Erdos-Renyi graphs G(n_i, p_i), n_i ~
    U{1000..200000}, p_i = 12 / (n_i - 1) (average degree 12), seed 4 (sizes and per-graph
    generator seeds drawn from numpy's default_rng(4)); one nkpr_run_batch to a certified L1
    error <= tol for every graph, mixed next to plain FP64.  Edge-updates count every arc of
    the union once per sweep."""
    from concurrent.futures import ThreadPoolExecutor

    import torch

    from networkit_amd import capi
    from networkit_amd import generators as gen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    G = args.graphs
    rng = np.random.default_rng(4)
    sizes = rng.integers(1000, 200001, size=G)
    gseeds = rng.integers(1, 1 << 62, size=G)
    t0 = time.time()

    def make(i):
        n = int(sizes[i])
        return gen.erdos_renyi(n, 12.0 / (n - 1), seed=int(gseeds[i]), directed=False)

    with ThreadPoolExecutor(16) as pool:
        edges = list(pool.map(make, range(G)))
    offs = np.zeros(G + 1, dtype=np.int64)
    offs[1:] = np.cumsum(sizes)
    z = int(offs[-1])
    # undirected in-lists of the union (edge (u, v) adds v to row u, then u to row v), sorted
    # stably by row on the GPU
    tu = torch.cat([torch.from_numpy((u + offs[i]).astype(np.int32)) for i, (u, _) in enumerate(edges)]).to(dev)
    tv = torch.cat([torch.from_numpy((v + offs[i]).astype(np.int32)) for i, (_, v) in enumerate(edges)]).to(dev)
    del edges
    rows = torch.stack([tu, tv], dim=1).reshape(-1)
    srcs = torch.stack([tv, tu], dim=1).reshape(-1)
    del tu, tv
    order = torch.argsort(rows, stable=True)
    colptr = torch.zeros(z + 1, dtype=torch.int64, device=dev)
    colptr[1:] = torch.cumsum(torch.bincount(rows, minlength=z), 0)
    src = srcs[order].contiguous()
    del rows, srcs, order
    m = int(colptr[-1])
    torch.cuda.synchronize()
    log("%d Erdos-Renyi graphs: z=%d m=%d arcs, built in %.1f s" % (G, z, m, time.time() - t0))
    t0 = time.time()
    dg = capi.DeviceGraph(z, colptr, src, device=True, graph_offsets=offs.astype(np.uint64))
    del colptr, src
    torch.cuda.synchronize()
    log("batch handle in %.1f s" % (time.time() - t0))
    stream = torch.cuda.ExternalStream(dg.stream())
    x = torch.empty(z, dtype=torch.float64, device=dev)
    start = torch.repeat_interleave(torch.from_numpy(1.0 / sizes.astype(np.float64)).to(dev),
                                    torch.from_numpy(sizes).to(dev))
    seg = torch.repeat_interleave(torch.arange(G, device=dev), torch.from_numpy(sizes).to(dev))

    def solve(precision, tol):
        torch.cuda.current_stream().synchronize()
        with torch.cuda.stream(stream):
            x.copy_(start)
        p = capi.default_params(scores_on_device=1, precision=precision, tol=tol,
                                stop=capi.NKPR_STOP_L1_ERROR, timing=1)
        st, its, mx = dg.run_batch(x, params=p)
        return st

    # per-graph sweep counts of the reference rule (FP64, L2 step <= 1e-8; not timed)
    torch.cuda.current_stream().synchronize()
    with torch.cuda.stream(stream):
        x.copy_(start)
    _, its_rule, _ = dg.run_batch(x, params=capi.default_params(
        scores_on_device=1, precision=capi.NKPR_FP64, tol=1e-8))
    solve(capi.NKPR_FP64, 1e-13)
    torch.cuda.synchronize()
    ref = x.clone()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_graphs_baseline(sizes, gseeds, its_rule, args.cpu_seconds)

    def timed(precision):
        for _ in range(args.warmup):
            solve(precision, args.tol)
        torch.cuda.synchronize()
        t = time.perf_counter()
        sts = [solve(precision, args.tol) for _ in range(args.steps)]
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t) / args.steps
        err = torch.zeros(G, dtype=torch.float64, device=dev).index_add_(0, seg, (x - ref).abs())
        return sec, sts, float(err.max())

    sec_m, st_m, err_m = timed(capi.NKPR_MIXED)
    sec_d, st_d, err_d = timed(capi.NKPR_FP64)
    s0, d0 = st_m[0], st_d[0]
    traffic, traffic_src = measured_traffic(m, z, "graphs")
    b32 = algorithmic_bytes(m, z, False, "fp32")
    b64 = algorithmic_bytes(m, z, False, "fp64")
    ms32 = sum(st["ms_sweeps_fp32"] for st in st_m) / max(1, sum(st["sweeps_fp32"] for st in st_m))
    ms64 = sum(st["ms_sweeps_fp64"] for st in st_d) / max(1, sum(st["sweeps_fp64"] for st in st_d))
    ach32 = b32 / (ms32 / 1e3) / 1e9
    ach64 = b64 / (ms64 / 1e3) / 1e9
    log("mixed: %.1f ms, %d fp32 + %d fp64 sweeps, %d refinements, max L1 err %.2e, bound %.2e; "
        "fp32 sweep %.3f ms = %.0f GB/s" % (sec_m * 1e3, s0["sweeps_fp32"], s0["sweeps_fp64"],
                                             s0["refinements"], err_m, s0["l1_error_bound"], ms32, ach32))
    log("fp64: %.1f ms, %d sweeps, max L1 err %.2e; fp64 sweep %.3f ms = %.0f GB/s" % (
        sec_d * 1e3, d0["sweeps_fp64"], err_d, ms64, ach64))
    sweeps_m = s0["sweeps_fp32"] + s0["sweeps_fp64"]
    out = {
        "metric": "PageRank edge-updates/s on a batch of graphs, mixed vs FP64 at equal L1 error",
        "value": m * sweeps_m / sec_m, "unit": "edge-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec_m * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": "one batched PageRank run over %d undirected Erdos-Renyi graphs, "
                               "n ~ U{1000..200000}, average degree 12, seed 4, certified L1 "
                               "error <= %g per graph (damp 0.85, no sink handling)"
                               % (G, args.tol), "graphs": G, "z": z, "m": m},
        "roofline": {"bound": "hbm", "achieved": ach32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach32 / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "FP32 GMRES operator: " + operator_name(m),
                     "bytes_per_sweep": b32, "ms_per_sweep": ms32},
        "cpu_baseline": cpu,
        "mixed": {"ms_per_solve": sec_m * 1e3, "sweeps_fp32": s0["sweeps_fp32"],
                  "sweeps_fp64": s0["sweeps_fp64"], "refinements": s0["refinements"],
                  "max_l1_error": err_m, "l1_error_bound": s0["l1_error_bound"]},
        "fp64": {"ms_per_solve": sec_d * 1e3, "sweeps": d0["sweeps_fp64"], "max_l1_error": err_d,
                 "ms_per_sweep": ms64, "achieved_GBs": ach64},
        "speedup_vs_fp64": sec_d / sec_m,
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """Differentiable PageRank: directed R-MAT scale 20, FP32 weights U(0.5,1.5), loss
    L = sum_i c_i x_i with c ~ N(0,1) (seed 5).  Forward = solve to a certified L1 error <= tol,
    backward = nkpr_adjoint (dL/dw for every edge and dL/dt) to a relative residual <= tol, both
    in mixed precision; the FP32-only variant (restarted FP32 GMRES adjoint, FP32 power
    forward at tol 1e-5) is reported beside it."""
    import torch

    from networkit_amd import autograd as nkad
    from networkit_amd import capi
    from networkit_amd import generators as gen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = gen.rmat(20, 16, seed=5, device="cuda")
    torch.cuda.synchronize()
    z = g["z"]
    m = g["src"].numel()
    rng = np.random.default_rng(5)
    c = torch.tensor(rng.normal(size=z), device=dev)
    res = {}
    for name, prec, tol in (("mixed", capi.NKPR_MIXED, args.tol), ("fp32", capi.NKPR_FP32, 1e-5)):
        op = nkad.PageRankOperator(g["colptr"], g["src"], tol=tol, precision=prec)
        w = g["weight"].clone().requires_grad_(True)
        fwd, bwd = [], []
        for it in range(args.warmup + args.steps):
            w.grad = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x = op(w)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            L = (x.double() * c).sum()
            L.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if it >= args.warmup:
                fwd.append(t1 - t0)
                bwd.append(t2 - t1)
        f = sum(fwd) / len(fwd)
        b = sum(bwd) / len(bwd)
        res[name] = {"forward_ms": f * 1e3, "backward_ms": b * 1e3, "backward_over_forward": b / f,
                     "forward_sweeps": op.last_forward["iterations"],
                     "backward_sweeps": op.last_backward["iterations"],
                     "backward_residual": op.last_backward["l1_error_bound"]}
        log("%s: forward %.2f ms (%d sweeps), backward %.2f ms (%d sweeps), ratio %.2f" % (
            name, f * 1e3, res[name]["forward_sweeps"], b * 1e3, res[name]["backward_sweeps"],
            b / f))
    r = res["mixed"]
    cpu = None
    if not args.no_cpu:
        import oracle
        og = host_in_lists(g["colptr"], g["src"], g["weight"])
        threads = cpu_threads()
        t = time.perf_counter()
        _, it, _ = oracle.pagerank(og, threads=threads)
        tf = time.perf_counter() - t
        log("cpu baseline: forward run() %.1f ms (%d sweeps) on %d threads" % (tf * 1e3, it,
                                                                              threads))
        cpu = {"value": tf * 1e3, "unit": "ms per forward PageRank::run", "cores": threads,
               "kind": "port",
               "sample": "the oracle's PageRank::run on the same graph and weights to the "
                         "reference rule (L2 step <= 1e-8, %d sweeps) on %d host threads; the "
                         "reference has no backward pass" % (it, threads),
               "host": host_info()}
    out = {"metric": "differentiable PageRank forward + backward time, backward / forward",
           "value": r["backward_over_forward"], "unit": "ratio", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": r["forward_ms"] + r["backward_ms"], "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64", "data": "synthetic",
           "config": {"workload": "forward and adjoint of PageRank on directed R-MAT scale 20, "
                                  "edge factor 16, FP32 weights U(0.5,1.5), loss sum c_i x_i, "
                                  "c ~ N(0,1), seed 5; tol %g" % args.tol, "n": z, "m": m},
           "mixed": r, "fp32": res["fp32"], "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_algebraic(args):
    """AlgebraicPageRank (AlgebraicPageRank.hpp:94-119) on NetworKit's CSR matrix of the
    benchmark's R-MAT graph: M = d A^T D^-1 as CSRGeneralMatrix<double> (uint64 row offsets and
    column indices, FP64 values), resident in HBM.  One step = one run from 1/n to
    ||y - x||_2 <= tol (the reference's rule), then y / sum(y), through the handle the C++
    drop-in keeps (nkpr_algebraic_create once, outside the timed region, like the reference's
    constructor building M; nkpr_algebraic_run per step).  Its dominant kernel is the FP64 sweep
    of the handle (binned at this size), timed with HIP events on the handle's stream.  The
    stateless nkpr_algebraic_pagerank (CSR kernels on M as given) is timed beside it."""
    import torch

    from networkit_amd import capi
    from networkit_amd import generators as gen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.time()
    g = gen.rmat(args.scale, args.edge_factor, seed=2, device="cuda")
    z = g["z"]
    colptr, src = g["colptr"], g["src"].to(torch.int64)
    w = g["weight"].to(torch.float64)
    # M(v, u) = d w(u, v) / outdeg_w(u): the rows are the in-lists, the columns their sources
    deg = torch.zeros(z, dtype=torch.float64, device=dev).index_add_(0, src, w)
    vals = 0.85 * w / deg[src]
    del w, deg
    mat = capi.DeviceCSR(z, z, colptr, src, vals)
    nnz = mat.nnz
    del g, vals
    torch.cuda.synchronize()
    log("algebraic: M with %d rows, %d entries, built in %.1f s" % (z, nnz, time.time() - t0))
    scores = torch.empty(z, dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    h = capi.DeviceAlgebraic(mat)
    prep_ms = (time.perf_counter() - t0) * 1e3
    its = 0
    t0 = time.perf_counter()
    _, its, st = h.run(tol=args.tol, scores=scores, stats=True)  # builds the binned copy
    first_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(max(args.warmup - 1, 0)):
        h.run(tol=args.tol, scores=scores)
    torch.cuda.synchronize()
    stats = []
    t = time.perf_counter()
    for _ in range(args.steps):
        _, its, st = h.run(tol=args.tol, scores=scores, stats=True)
        stats.append(st)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t) / args.steps
    ms_sweep = sum(x["ms_sweeps_fp64"] for x in stats) / sum(x["sweeps_fp64"] for x in stats)
    bytes_step = algorithmic_bytes(nnz, z, True, "fp64")
    ach = bytes_step / (ms_sweep / 1e3) / 1e9
    log("algebraic handle: built in %.0f ms, first run %.0f ms; %.1f ms per run (%d iterations), "
        "%.3f ms per sweep = %.0f GB/s" % (prep_ms, first_ms, sec * 1e3, its, ms_sweep, ach))
    ref_scores = scores.clone()
    # the stateless CSR-kernel run on the same matrix, and its product step alone (HIP events)
    t = time.perf_counter()
    _, its_csr = mat.algebraic_pagerank(tol=args.tol, scores=scores)
    torch.cuda.synchronize()
    sec_csr = time.perf_counter() - t
    l1_csr = float((scores - ref_scores).abs().sum())
    x = torch.full((z,), 1.0 / z, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    sq = torch.empty(1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    reps = 10
    mat.pagerank_step(x, 0.15 / z, y, sq, stream=stream.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(reps):
        a, b = (x, y) if i % 2 == 0 else (y, x)
        mat.pagerank_step(a, 0.15 / z, b, sq, stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_step = e0.elapsed_time(e1) / reps
    log("algebraic CSR kernels: %.1f ms per run (%d iterations), product step %.3f ms = %.0f GB/s "
        "of 24 nnz + 24 n; L1 between the two runs %.2e" % (
            sec_csr * 1e3, its_csr, ms_step, (24 * nnz + 24 * z) / (ms_step / 1e3) / 1e9, l1_csr))
    cpu = None
    if not args.no_cpu:
        import oracle
        hm = oracle.CSR(z, z, mat.row_idx.cpu().numpy().view(np.uint64),
                        mat.col_idx.cpu().numpy().view(np.uint64), mat.values.cpu().numpy())
        threads = int(oracle.num_threads())
        t = time.perf_counter()
        oracle.algebraic_pagerank(hm, tol=0.0, max_iterations=1)
        t1 = time.perf_counter() - t
        k = int(max(2, min(40, args.cpu_seconds / max(t1, 1e-3))))
        t = time.perf_counter()
        oracle.algebraic_pagerank(hm, tol=0.0, max_iterations=k)
        tk = time.perf_counter() - t
        per = max((tk - t1) / (k - 1), 1e-9)
        log("cpu baseline: run() with 1 iteration %.2f s, with %d %.2f s on %d threads" % (
            t1, k, tk, threads))
        cpu = {"value": nnz * k / tk, "unit": "edge-updates/s", "cores": threads,
               "kind": "port",
               "sample": "the oracle's AlgebraicPageRank::run on the same matrix capped at %d "
                         "iterations (%.2f s) and at 1 (%.2f s) on %d host threads" % (
                             k, tk, t1, threads),
               "ms_per_iteration": per * 1e3,
               "run_to_reference_rule_s_extrapolated": max(t1 - per, 0.0) + its * per,
               "host": host_info()}
    out = {"metric": "AlgebraicPageRank edge-updates/s (NetworKit CSR, FP64, L2 step <= tol)",
           "value": nnz * its / sec, "unit": "edge-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
           "data": "synthetic",
           "config": {"workload": "AlgebraicPageRank run (damp 0.85, tol %g, L2 step) on the "
                                  "CSR matrix d A^T D^-1 of a directed R-MAT graph, scale %d, "
                                  "edge factor %d, a/b/c/d=.57/.19/.19/.05, duplicates and "
                                  "self-loops removed, FP32 weights U(0.5,1.5) as FP64, seed 2"
                                  % (args.tol, args.scale, args.edge_factor),
                      "n": z, "nnz": nnz, "iterations": its, "handle_build_ms": prep_ms,
                      "first_run_ms": first_ms},
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "FP64 sweep of the handle built from M (binned: "
                                  "bin_scatter64_kernel + bin_gather64_kernel + "
                                  "bin_reduce64_kernel), 20 nnz + 40 n bytes",
                        "bytes_per_sweep": bytes_step, "ms_per_sweep": ms_sweep},
           "csr_kernels": {"ms_per_run": sec_csr * 1e3, "iterations": its_csr,
                           "ms_per_step": ms_step, "l1_vs_handle": l1_csr},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_apr(args):
    """ApproximatePageRank (ApproximatePageRank.cpp:18-61) on the undirected random hyperbolic
    graph (n = rhg_nodes, average degree 20, exponent 3, seed 3): residual pushes from one seed
    (the node of largest degree among ids 0..999), alpha 0.1, epsilon apr_eps.  One step = one
    run to the stopping rule (every residual below eps times the weighted degree); the metric
    is runs per second, the device's push counts are reported beside it."""
    import torch

    from networkit_amd import capi
    from networkit_amd import generators as gen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = gen.rhg(args.rhg_nodes, 20.0, 3.0, seed=3)
    z = g["z"]
    colptr, src = g["colptr"], g["src"]
    deg = (colptr[1:1001] - colptr[:1000]).cpu().numpy()
    seed = int(np.argmax(deg))
    alpha, eps = 0.1, args.apr_eps
    apr = capi.DeviceAPR(z, colptr, src, None, device=True)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        apr.run([seed], alpha, eps)
    t = time.perf_counter()
    for _ in range(args.steps):
        ids, p, r, st = apr.run([seed], alpha, eps)
    sec = (time.perf_counter() - t) / args.steps
    mass = float(p.sum() + r.sum())
    log("apr: %.2f ms per run, %d rounds, %d pushes, %d edge pushes, %d touched, mass %.15f" % (
        sec * 1e3, st["rounds"], st["pushes"], st["edge_pushes"], st["touched"], mass))
    # per push: id 4 + vol 8 of the node, residual exchange 8, p 8; per edge: id 4, vol 8 of the
    # neighbour, the 8-byte atomic (read and write)
    bytes_run = 28 * st["pushes"] + 28 * st["edge_pushes"]
    ach = bytes_run / sec / 1e9
    cpu = None
    if not args.no_cpu:
        import oracle
        host = host_in_lists(colptr, src)
        runs, spent = 0, 0.0
        while runs == 0 or (spent < args.cpu_seconds and runs < 20):
            t0 = time.perf_counter()
            po, ro, _ = oracle.approximate_pagerank(z, host.in_ptr, host.in_src, None, [seed],
                                                    alpha, eps)
            spent += time.perf_counter() - t0
            runs += 1
        tc = spent / runs
        dev_l1 = float(np.abs(po[ids] - p).sum() + po.sum() - po[ids].sum())
        log("cpu baseline: oracle FIFO run %.3f s (%d runs), touched %d, |p - p_oracle|_1 %.2e"
            % (tc, runs, int(((po > 0) | (ro > 0)).sum()), dev_l1))
        cpu = {"value": 1.0 / tc, "unit": "runs/s", "cores": 1, "kind": "port",
               "sample": "%d whole runs of the oracle's ApproximatePageRank::run (the "
                         "reference's FIFO queue, one thread, as the reference runs) from the "
                         "same seed: %.3f s each" % (runs, tc),
               "host": host_info()}
    out = {"metric": "ApproximatePageRank runs/s", "value": 1.0 / sec,
           "unit": "runs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
           "config": {"workload": "ApproximatePageRank from one seed (alpha 0.1, epsilon %g) "
                                  "on an undirected threshold random hyperbolic graph, n=%d, "
                                  "average degree 20, exponent 3, seed 3"
                                  % (eps, args.rhg_nodes),
                      "n": z, "m": int(src.numel()), "seed": seed, "rounds": st["rounds"],
                      "pushes": st["pushes"], "edge_pushes": st["edge_pushes"],
                      "touched": st["touched"]},
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "whole run (push rounds, frontier swaps, compaction)",
                        "bytes_per_sweep": bytes_run, "ms_per_sweep": sec * 1e3},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_power(args):
    """EigenvectorCentrality (EigenvectorCentrality.cpp:21-66, tol 1e-9, its default) and
    KatzCentrality (KatzCentrality.cpp:28-76: alpha = 1 / (1 + max out-degree), beta 0.1, tol
    1e-8, in-edges, its defaults) on the benchmark's R-MAT graph, through nkpr_power_run on a
    handle with unit out-degrees (its coefficients are the weights).  One step = one
    eigenvector run and one Katz run from the reference's start vectors.  Metric: edge-updates
    per second over both (m x sweeps / time).  The dominant kernel is the FP64 sweep (binned
    at this size), timed with HIP events on the library's stream."""
    import torch

    from networkit_amd import capi
    from networkit_amd import generators as gen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = gen.rmat(args.scale, args.edge_factor, seed=2, device="cuda")
    z = g["z"]
    colptr, src, w = g["colptr"], g["src"], g["weight"]
    m = src.numel()
    ones = torch.ones(z, dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    dg = capi.DeviceGraph(z, colptr, src, w, out_wdeg=ones, device=True)
    prep = time.perf_counter() - t0
    outdeg = torch.bincount(src.to(torch.int64), minlength=z)
    alpha = 1.0 / (1.0 + float(outdeg.max().item()))
    del outdeg
    x = torch.empty(z, dtype=torch.float64, device=dev)

    def runs():
        se = dg.power_run(x, capi.default_power_params(kind=capi.NKPR_EIGENVECTOR, tol=1e-9,
                                                       scores_on_device=1))
        sk = dg.power_run(x, capi.default_power_params(kind=capi.NKPR_KATZ, alpha=alpha,
                                                       beta=0.1, tol=1e-8, scores_on_device=1))
        return se, sk

    for _ in range(args.warmup):
        runs()
    torch.cuda.synchronize()
    t = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        stats.append(runs())
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t) / args.steps
    se, sk = stats[-1]
    sweeps = se["sweeps_fp64"] + sk["sweeps_fp64"]
    ms_sweep = sum(a["ms_sweeps_fp64"] + b["ms_sweeps_fp64"] for a, b in stats) / \
        sum(a["sweeps_fp64"] + b["sweeps_fp64"] for a, b in stats)
    bytes_sweep = algorithmic_bytes(m, z, True, "fp64")
    ach = bytes_sweep / (ms_sweep / 1e3) / 1e9
    log("power: eigenvector %d sweeps, Katz %d sweeps (alpha %.3g); %.1f ms per step, %.3f ms per "
        "sweep = %.0f GB/s; handle built in %.2f s" % (se["sweeps_fp64"], sk["sweeps_fp64"], alpha,
                                                        sec * 1e3, ms_sweep, ach, prep))
    cpu = None
    if not args.no_cpu:
        import oracle
        og = host_in_lists(colptr, src, w)
        threads = int(oracle.num_threads())
        t = time.perf_counter()
        oracle.katz(og, alpha=alpha, beta=0.1, tol=0.0, max_iterations=1)
        t1 = time.perf_counter() - t
        k = int(max(2, min(40, args.cpu_seconds / max(t1, 1e-3))))
        t = time.perf_counter()
        oracle.katz(og, alpha=alpha, beta=0.1, tol=0.0, max_iterations=k)
        tk = time.perf_counter() - t
        per = max((tk - t1) / (k - 1), 1e-9)
        est = max(t1 - per, 0.0) + sweeps * per
        log("cpu baseline: Katz run() capped at 1 sweep %.2f s, at %d sweeps %.2f s on %d threads; "
            "%.3f s per sweep" % (t1, k, tk, threads, per))
        cpu = {"value": m * sweeps / est, "unit": "edge-updates/s", "cores": threads,
               "kind": "port",
               "sample": "the oracle's KatzCentrality::run (the restatement of "
                         "KatzCentrality.cpp:28-76; the eigenvector loop is the same sweep) on "
                         "the same graph capped at %d sweeps (%.2f s) and at 1 (%.2f s) on %d "
                         "host threads; value = m x %d sweeps (the GPU step's) / (set-up + %d x "
                         "the time per sweep)" % (k, tk, t1, threads, sweeps, sweeps),
               "ms_per_sweep": per * 1e3, "host": host_info()}
    out = {"metric": "Eigenvector + Katz centrality edge-updates/s (FP64, reference stop rules)",
           "value": m * sweeps / sec, "unit": "edge-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
           "data": "synthetic",
           "config": {"workload": "EigenvectorCentrality (tol 1e-9) and KatzCentrality (alpha = "
                                  "1/(1 + max out-degree), beta 0.1, tol 1e-8) on a directed "
                                  "R-MAT graph, scale %d, edge factor %d, a/b/c/d=.57/.19/.19/.05,"
                                  " duplicates and self-loops removed, FP32 weights U(0.5,1.5), "
                                  "seed 2" % (args.scale, args.edge_factor),
                      "n": z, "m": m, "eigenvector_sweeps": se["sweeps_fp64"],
                      "katz_sweeps": sk["sweeps_fp64"], "katz_alpha": alpha},
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "FP64 sweep of the unit-out-degree handle (binned: "
                                  "bin_scatter64_kernel + bin_gather64_kernel + "
                                  "bin_reduce64_kernel), 20 m + 40 n bytes",
                        "bytes_per_sweep": bytes_sweep, "ms_per_sweep": ms_sweep},
           "cpu_baseline": cpu}
    dg.close()
    print(json.dumps(out), flush=True)


def write_rmat_nkbg(path, scale, chunks=32):
    """The benchmark's R-MAT graph (seed 2, FP32 weights) at `scale` written as nkbg002 by the
    library's writer (out-lists and in-lists in storage order, as NetworKit's writer takes
    them from a Graph).  Returns (nodes, arcs)."""
    import torch

    from networkit_amd import capi
    from networkit_amd import generators as gen
    g = gen.rmat(scale, 16, seed=2, device="cuda")
    z = g["z"]
    colptr, src, w = g["colptr"], g["src"], g["weight"]
    m = src.numel()
    dst = torch.repeat_interleave(torch.arange(z, device=src.device, dtype=torch.int32),
                                  (colptr[1:] - colptr[:-1]).to(torch.int64))
    order = torch.argsort(src, stable=True)
    out_ptr = torch.zeros(z + 1, dtype=torch.int64, device=src.device)
    out_ptr[1:] = torch.cumsum(torch.bincount(src.to(torch.int64), minlength=z), 0)
    host = lambda t: t.cpu().numpy()
    w64 = w.to(torch.float64)
    capi.write_nkbg(path, z, True, host(out_ptr).astype(np.uint64), host(dst[order]),
                    host(w64[order]), host(colptr).astype(np.uint64), host(src), host(w64),
                    None, chunks=chunks)
    del g, colptr, src, w, dst, order, out_ptr, w64
    torch.cuda.synchronize()
    return z, m


def bench_nkbg(args):
    """NetworKit binary graph files (NetworkitBinaryReader.cpp:19-237) to device lists: the
    R-MAT graph (scale nkbg_scale, edge factor 16, seed 2, FP32 weights) written by the
    library's writer with the reference writer's default of 32 chunks, then read straight into
    the flat pull lists PageRank takes (no per-node vectors) and copied to HBM.  One step =
    nkpr_nkbg_read (the file memory-mapped and decoded on every host thread, one chunk per
    thread at a time) + the copy of colptr, src and the weights to device memory.  Metric: arcs
    per second, file to device."""
    import ctypes
    import torch

    from networkit_amd import capi

    torch.cuda.set_device(0)
    path = os.path.join(args.nkbg_dir, "nkpr_bench_rmat%d.nkbg" % args.nkbg_scale)
    t0 = time.perf_counter()
    z, m = write_rmat_nkbg(path, args.nkbg_scale)
    size = os.path.getsize(path)
    log("nkbg: R-MAT scale %d (%d arcs) written in %.1f s, %.2f GB" % (
        args.nkbg_scale, m, time.perf_counter() - t0, size / 1e9))
    L = capi.lib()
    dev = torch.device("cuda:0")

    def step():
        h = capi.Nkbg()
        t = time.perf_counter()
        capi.check(L.nkpr_nkbg_read(os.fsencode(path), 0, ctypes.byref(h)))
        t_read = time.perf_counter() - t
        try:
            n = int(h.nodes)
            mm = int(np.ctypeslib.as_array(ctypes.cast(h.colptr, ctypes.POINTER(ctypes.c_uint64)),
                                           (n + 1,))[-1])
            arrays = [(h.colptr, ctypes.c_int64, n + 1), (h.src, ctypes.c_int32, mm)]
            if h.weighted:
                arrays.append((h.weight, ctypes.c_double, mm))
            out = []
            for ptr, ct, cnt in arrays:
                a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), (cnt,))
                out.append(torch.from_numpy(a).to(dev, non_blocking=False))
            torch.cuda.synchronize()
        finally:
            L.nkpr_nkbg_free(ctypes.byref(h))
        return time.perf_counter() - t, t_read, out

    try:
        for _ in range(args.warmup):
            step()
        tot, rd = [], []
        for _ in range(args.steps):
            t_all, t_read, out = step()
            tot.append(t_all)
            rd.append(t_read)
        sec, sec_read = sum(tot) / len(tot), sum(rd) / len(rd)
        same = bool(int(out[0][-1].item()) == m and int(out[1].numel()) == m)
        del out
    finally:
        os.remove(path)
    log("nkbg: file to device %.3f s (%.3g arcs/s, %.2f GB/s of file): decode %.3f s on %d host "
        "threads, copies %.3f s" % (sec, m / sec, size / sec / 1e9, sec_read, os.cpu_count() or 1,
                                    sec - sec_read))
    cpu = None
    if not args.no_cpu:
        from oracle import nkbg as onkbg
        small = os.path.join(args.nkbg_dir, "nkpr_bench_rmat16.nkbg")
        zs, ms = write_rmat_nkbg(small, 16)
        data = open(small, "rb").read()
        os.remove(small)
        t = time.perf_counter()
        onkbg.read(data)
        tc = time.perf_counter() - t
        log("cpu baseline: the oracle's reader on R-MAT scale 16 (%d arcs): %.2f s" % (ms, tc))
        cpu = {"value": ms / tc, "unit": "arcs/s", "cores": 1, "kind": "port",
               "sample": "the oracle's restatement of NetworkitBinaryReader::read (Python, one "
                         "thread, into per-node lists as the reference builds them) on the same "
                         "generator's scale-16 graph (%d arcs, %.2f s)" % (ms, tc),
               "host": host_info()}
    out = {"metric": "NetworKit binary graph file to device pull lists, arcs/s",
           "value": m / sec, "unit": "arcs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
           "config": {"workload": "nkbg002 file (32 chunks, FP32 weights) of a directed R-MAT "
                                  "graph, scale %d, edge factor 16, seed 2, read into flat "
                                  "in-lists and weights and copied to device memory"
                                  % args.nkbg_scale,
                      "n": z, "m": m, "file_bytes": size, "decode_s": sec_read,
                      "copy_s": sec - sec_read, "lists_complete": same},
           "roofline": {"bound": "hbm", "achieved": size / sec / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": size / sec / 1e9 / HBM_PEAK_GBS,
                        "traffic": None,
                        "kernel": "none: the host decode of the file bounds the path (file "
                                  "bytes per second, file to device)",
                        "bytes_per_sweep": size, "ms_per_sweep": sec * 1e3},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def operator_name(m):
    """The kernels of one application of the GMRES operator: the library takes the binned
    two-pass operator for graphs of 2^26 edges and more (NKPR_BINNED=0/1 forces either)."""
    env = os.environ.get("NKPR_BINNED", "")
    binned = (int(env) != 0) if env else m >= 1 << 26
    return ("binned (bin_scatter_kernel + bin_gather_kernel + bin_reduce_kernel)" if binned
            else "pull sweep (sweep_kernel<float> + combine_kernel<float>)")


def measured_traffic(m, n, workload):
    """HBM bytes per application of the GMRES operator, read from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (scaled by the counters' reading of a
    1 GiB copy; not measured in this run).  Returns (bytes, source) with bytes None when no
    summary exists for this graph and operator."""
    name = {"pagerank": "pagerank_rmat26_traffic.json",
            "rhg": "pagerank_rhg10m_traffic.json",
            "graphs": "pagerank_er512_batch_traffic.json"}.get(workload)
    if name is None:
        return None, None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if t.get("m") != m or t.get("n") != n or t.get("operator") != operator_name(m):
        return None, None
    return t["operator_traffic_bytes"], {"source": "profiles/" + name, "measured_in_run": False}


def host_info():
    """The host the CPU baseline ran on: logical CPUs, the process's CPU affinity and lscpu's
    model, sockets, cores per socket and threads per core."""
    import subprocess
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                             text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        out = ""
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            info[keys[k.strip()]] = v.strip()
    return info


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))


def host_in_lists(colptr, src, weight=None):
    """NetworKit element types on the host (uint64 offsets and ids, FP64 weights) from the
    device CSC of a workload."""
    import oracle
    cp = colptr.cpu().numpy() if hasattr(colptr, "cpu") else np.asarray(colptr)
    sr = src.cpu().numpy() if hasattr(src, "cpu") else np.asarray(src)
    w = None
    if weight is not None:
        w = (weight.cpu().numpy() if hasattr(weight, "cpu") else np.asarray(weight)).astype(np.float64)
    return oracle.InLists(len(cp) - 1, cp.astype(np.uint64), sr.astype(np.uint64), w)


def cpu_run_baseline(og, m, seconds, its_ref, what):
    """The oracle's PageRank::run (a restatement of PageRank.cpp:20-104 in NetworKit's element
    types, guided OpenMP schedule) timed on the host cores on a bounded sample: run() capped at
    1 iteration, then at k iterations (k sized to ~`seconds` of work).  Their difference gives
    the per-sweep time and the set-up (the weighted out-degrees of PageRank.cpp:29-30); the
    time of run() to the reference rule (L2 step <= 1e-8, its_ref sweeps, the count the GPU's
    FP64 run with that rule took, which equals the oracle's) is extrapolated from them."""
    import oracle
    threads = cpu_threads()
    t = time.perf_counter()
    oracle.pagerank(og, tol=0.0, max_iterations=1, threads=threads)
    t1 = time.perf_counter() - t
    k = int(max(2, min(40, seconds / max(t1, 1e-3))))
    t = time.perf_counter()
    oracle.pagerank(og, tol=0.0, max_iterations=k, threads=threads)
    tk = time.perf_counter() - t
    per = max((tk - t1) / (k - 1), 1e-9)
    setup = max(t1 - per, 0.0)
    est = setup + its_ref * per
    log("cpu baseline: run() with 1 sweep %.2f s, with %d sweeps %.2f s on %d threads; "
        "%.3f s per sweep, set-up %.2f s" % (t1, k, tk, threads, per, setup))
    return {"value": m * its_ref / est, "unit": "edge-updates/s", "cores": threads,
            "kind": "port",
            "sample": "%s: the oracle's PageRank::run capped at %d iterations (%.2f s) and at 1 "
                      "(%.2f s) on %d host threads; value = m x %d sweeps / (set-up + %d x the "
                      "time per sweep), a whole run() to the reference rule" % (
                          what, k, tk, t1, threads, its_ref, its_ref),
            "ms_per_sweep": per * 1e3, "setup_ms": setup * 1e3,
            "reference_rule_sweeps": int(its_ref),
            "run_to_reference_rule_s_extrapolated": est, "host": host_info()}


def cpu_ppr_baseline(og, nn, m, seeds, its_rule, seconds):
    """The oracle's PageRank::run with each column's teleport vector (the personalised
    replacement of (1-d)/n; the reference has no teleport vector, so its loop is otherwise
    unchanged) to the reference rule, one column after another on the host cores, for the
    first columns until ~`seconds`; the whole batch is extrapolated from the rate per sweep and
    the per-column sweep counts of the GPU's FP64 run with that rule."""
    import oracle
    threads = cpu_threads()
    done, sweeps, t_all = 0, 0, 0.0
    for j, sd in enumerate(seeds):
        tv = np.zeros(nn)
        np.add.at(tv, np.asarray(sd, np.int64), 1.0)
        tv /= tv.sum()
        t = time.perf_counter()
        _, it, _ = oracle.pagerank(og, teleport=tv, threads=threads)
        t_all += time.perf_counter() - t
        done += 1
        sweeps += it
        if it != int(its_rule[j]):
            log("cpu baseline: column %d took %d sweeps, the GPU %d" % (j, it, int(its_rule[j])))
        if t_all >= seconds:
            break
    per_sweep = t_all / max(sweeps, 1)
    est = per_sweep * float(np.sum(its_rule))
    log("cpu baseline: %d columns, %d sweeps in %.2f s on %d threads; batch of %d columns "
        "~%.1f s" % (done, sweeps, t_all, threads, len(seeds), est))
    return {"value": m * sweeps / t_all, "unit": "column-edge-updates/s", "cores": threads,
            "kind": "port",
            "sample": "the oracle's PageRank::run with the teleport vectors of the first %d "
                      "columns, each to the reference rule (L2 step <= 1e-8), one after another "
                      "on %d host threads: %d sweeps in %.2f s" % (done, threads, sweeps, t_all),
            "batch_s_extrapolated": est, "host": host_info()}


def cpu_graphs_baseline(sizes, gseeds, its_rule, seconds):
    """The oracle's PageRank::run on the graphs of the batch one after another (each generated
    again from its seed) to the reference rule, until ~`seconds`; the whole batch is
    extrapolated from the rate and the per-graph sweep counts of the GPU's FP64 run."""
    import oracle
    from networkit_amd import generators as gen
    threads = cpu_threads()
    done, eu, t_all, m_all = 0, 0.0, 0.0, 0
    for i in range(len(sizes)):
        n = int(sizes[i])
        u, v = gen.erdos_renyi(n, 12.0 / (n - 1), seed=int(gseeds[i]), directed=False)
        og = oracle.in_lists(n, u, v, None, False)
        t = time.perf_counter()
        _, it, _ = oracle.pagerank(og, threads=threads)
        t_all += time.perf_counter() - t
        if it != int(its_rule[i]):
            log("cpu baseline: graph %d took %d sweeps, the GPU %d" % (i, it, int(its_rule[i])))
        eu += og.m * it
        m_all += og.m
        done += 1
        if t_all >= seconds:
            break
    rate = eu / t_all
    log("cpu baseline: %d graphs (%d arcs) in %.2f s on %d threads" % (done, m_all, t_all,
                                                                       threads))
    return {"value": rate, "unit": "edge-updates/s", "cores": threads, "kind": "port",
            "sample": "the oracle's PageRank::run on the first %d graphs of the batch, each to "
                      "the reference rule (L2 step <= 1e-8), one after another on %d host "
                      "threads: %.2f s" % (done, threads, t_all),
            "host": host_info()}


if __name__ == "__main__":
    main()
