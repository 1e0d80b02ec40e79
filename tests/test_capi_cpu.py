"""CPU-side checks of the C ABI library: it loads, exports every symbol include/nkpr.h
declares, and its host-only entry points behave; no device compute is attempted here."""
import ctypes
import os
import re

import numpy as np
import pytest

from networkit_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "nkpr.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nkpr_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_expected_symbols():
    assert declared_symbols() == sorted(capi.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = os.popen("nm -D --defined-only %s" % capi.LIB_PATH).read()
    for name in declared_symbols():
        assert re.search(r"\bT %s\b" % name, out), name


def test_struct_layouts_match_the_compiler():
    """Every ctypes mirror has the size and field offsets the C compiler gives the struct of
    nkpr.h (a small program built with gcc prints sizeof and offsetof)."""
    import shutil
    import subprocess
    import tempfile
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"nkpr_csc": capi.Csc, "nkpr_graph_info": capi.GraphInfo,
               "nkpr_params": capi.Params, "nkpr_stats": capi.Stats,
               "nkpr_batch_stats": capi.BatchStats, "nkpr_power_params": capi.PowerParams,
               "nkpr_csr": capi.CsrMatrix, "nkpr_nkbg": capi.Nkbg,
               "nkpr_apr_stats": capi.AprStats}
    lines = ["#include <stddef.h>", "#include <stdio.h>", "#include <nkpr.h>", "int main(void) {"]
    for c, py in mirrors.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (c, c))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (c, f, c, f))
    lines += ["return 0;", "}"]
    include = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        with open(src, "w") as fh:
            fh.write("\n".join(lines))
        subprocess.run([cc, "-std=c11", "-I", include, "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for c, py in mirrors.items():
        assert int(got[c]) == ctypes.sizeof(py), c
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (c, f)]) == getattr(py, f).offset, (c, f)


def test_struct_sizes_match_header_layout():
    # sizes of the C structs (x86-64 SysV): keep ctypes mirrors in sync with nkpr.h
    assert ctypes.sizeof(capi.Csc) == 8 * 7 + 8
    assert ctypes.sizeof(capi.Params) == 8 * 3 + 4 * 8 + 8 + 8
    assert ctypes.sizeof(capi.BatchStats) == 8 * 5 + 8 * 4 + 4 * 2
    assert ctypes.sizeof(capi.Stats) == 8 * 13
    assert ctypes.sizeof(capi.GraphInfo) == 8 * 7 + 8 + 8
    assert ctypes.sizeof(capi.PowerParams) == 4 * 2 + 8 * 3 + 8 + 8


def test_params_default_is_reference_default():
    p = capi.default_params()
    assert p.damp == 0.85 and p.tol == 1e-8
    assert p.max_iterations == 2 ** 64 - 1
    assert p.norm == capi.NKPR_L2
    assert p.sink_handling == capi.NKPR_SINK_NONE
    assert p.normalized == 0 and p.precision == capi.NKPR_FP64
    assert p.stop == capi.NKPR_STOP_STEP
    assert capi.lib().nkpr_abi_version() == capi.ABI_VERSION == 3


def test_graph_create_without_device_reports_nodevice():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(capi.NkprError) as e:
        capi.DeviceGraph(3, np.array([0, 1, 2, 3]), np.array([1, 2, 0]))
    assert e.value.code == capi.NKPR_ERR_NODEVICE
    assert "device" in capi.lib().nkpr_last_error().decode()


def test_null_arguments_are_errors_not_crashes():
    L = capi.lib()
    assert L.nkpr_graph_create(None, None) == capi.NKPR_ERR_INVALID or \
        L.nkpr_graph_create(None, None) == capi.NKPR_ERR_NODEVICE
    assert L.nkpr_run(None, None, None, None) != capi.NKPR_OK
    assert L.nkpr_ppr_batch(None, None, 0, None, None, None, None, None, None, None) != capi.NKPR_OK
    assert L.nkpr_run_batch(None, None, None, None, None, None) != capi.NKPR_OK
    assert L.nkpr_adjoint(None, None, None, None, 1.0, None, None, None, None) != capi.NKPR_OK
    assert L.nkpr_graph_set_weights(None, None, None, None, 0) != capi.NKPR_OK
    assert L.nkpr_graph_create_batch(None, 0, None, None) == capi.NKPR_ERR_INVALID
    assert L.nkpr_graph_destroy(None) == capi.NKPR_OK
    assert L.nkpr_params_default(None) == capi.NKPR_ERR_INVALID
    assert L.nkpr_power_params_default(None) == capi.NKPR_ERR_INVALID
    assert L.nkpr_power_run(None, None, None, None) != capi.NKPR_OK


def test_power_params_default_is_reference_default():
    p = capi.default_power_params()
    assert p.kind == capi.NKPR_EIGENVECTOR and p.tol == 1e-8
    assert p.alpha == 0.0 and p.beta == 0.1  # KatzCentrality(G, 0, 0.1, 1e-8)
    assert p.max_iterations == 2 ** 64 - 1 and p.scores_on_device == 0


def test_ranking_order_score_desc_id_asc():
    rng = np.random.default_rng(1)
    s = rng.integers(0, 10, 1000) / 3.0
    ex = (rng.random(1000) > 0.05).astype(np.uint8)
    ids, vals = capi.ranking(s, ex)
    ref = sorted([i for i in range(1000) if ex[i]], key=lambda i: (-s[i], i))
    assert ids.tolist() == ref
    assert np.array_equal(vals, s[ids])


def test_ranking_with_nan_scores_is_a_total_order():
    # NaN scores rank after every number, ties among them by id; the order of the numbers is
    # the reference's (score descending, id ascending)
    rng = np.random.default_rng(2)
    s = rng.integers(0, 6, 5000) / 7.0
    s[rng.random(5000) < 0.2] = np.nan
    s[rng.random(5000) < 0.05] = np.inf
    s[rng.random(5000) < 0.05] = -np.inf
    ids, vals = capi.ranking(s)
    num = [i for i in range(5000) if not np.isnan(s[i])]
    nan = [i for i in range(5000) if np.isnan(s[i])]
    assert ids.tolist() == sorted(num, key=lambda i: (-s[i], i)) + nan
    assert len(ids) == 5000
