"""Hot sources of the binned FP32 operator (NKPR_BINNED_HOT; off by default): the edges of the
first `hot` relabelled sources skip pass 1 and the contribution array, and pass 2 gathers their
values from an LDS table.
Every contribution is computed as pass 1 computes it and added into the same integer
accumulators, so the operator is bit-identical with and without the bypass: on handles with
the same coefficients (out-degrees given, not summed on the device) mixed solves agree bit for
bit and take the same sweeps, and they certify against the converged oracle.  Cases: weighted
and unweighted graphs, both sink rules, 3000 and 16384 hot sources, and small work items
(NKPR_BINNED_ITEM=256) so that pass-2 items cut bins between contributions and hot entries."""
import numpy as np
import pytest

import oracle
from graphs import skewed_graph
from networkit_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("weighted", [True, False])
@pytest.mark.parametrize("item", ["256", None])
@pytest.mark.parametrize("sinks", [0, 1])
def test_hot_sources_bit_identical(monkeypatch, capfd, weighted, item, sinks):
    monkeypatch.setenv("NKPR_BINNED", "1")
    monkeypatch.setenv("NKPR_DEBUG", "1")
    if item:
        monkeypatch.setenv("NKPR_BINNED_ITEM", item)
    else:
        monkeypatch.delenv("NKPR_BINNED_ITEM", raising=False)
    z, u, v, w = skewed_graph(150000, 1200000, 71, weighted=weighted)
    g = oracle.in_lists(z, u, v, w, True)
    deg, _ = oracle.out_degrees(g)
    ref, _, _ = oracle.pagerank(g, tol=1e-15, norm=oracle.L1, max_iterations=5000,
                                sink_mode=sinks)
    w32 = None if g.in_w is None else g.in_w.astype(np.float32)
    capfd.readouterr()
    res = {}
    for hot in ("0", "3000", "16384", "auto"):
        monkeypatch.setenv("NKPR_BINNED_HOT", hot)
        dg = capi.DeviceGraph(z, g.in_ptr, g.in_src.astype(np.int32), w32, weight64=g.in_w,
                              out_wdeg=deg)
        x = np.full(z, 1.0 / z)
        st = dg.run(x, capi.default_params(precision=capi.NKPR_MIXED, tol=1e-9,
                                           stop=capi.NKPR_STOP_L1_ERROR, sink_handling=sinks))
        dg.close()
        err = capfd.readouterr().err
        assert "binned FP32 layout" in err, err[-1500:]
        want = {"auto": "16384"}.get(hot, hot)  # the skewed graph's top sources hold > 5%
        assert ("(%s hot sources)" % want) in err, err[-1500:]
        assert st["l1_error_bound"] <= 1e-9
        assert np.abs(x - ref).sum() <= 1e-9
        res[hot] = (x, st)
    for hot in ("3000", "16384", "auto"):
        assert np.array_equal(res[hot][0], res["0"][0]), hot
        for key in ("refinements", "sweeps_fp32", "sweeps_fp64"):
            assert res[hot][1][key] == res["0"][1][key]
    print("weighted %s item %s sinks %d: %d FP32 applications, L1 %.2e" % (
        weighted, item, sinks, res["0"][1]["sweeps_fp32"], np.abs(res["0"][0] - ref).sum()),
        flush=True)
