"""ctypes binding of the C ABI in include/nkpr.h (libnkpr.so).

This is the thinnest Python route to the device library: tests use it to call every entry
point directly, and the benchmark uses it to hand device-resident arrays (torch tensors) to
nkpr_graph_create / nkpr_run without a host round trip.  The Cython module `_core` is the
NetworKit-style API on top of the C++ drop-in; both end in the same library.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnkpr.so")

NKPR_OK = 0
NKPR_ERR_INVALID = 1
NKPR_ERR_HIP = 2
NKPR_ERR_NOMEM = 3
NKPR_ERR_INTERRUPTED = 4
NKPR_ERR_NODEVICE = 5
NKPR_L1, NKPR_L2 = 0, 1
NKPR_SINK_NONE, NKPR_SINK_DISTRIBUTE = 0, 1
NKPR_FP64, NKPR_MIXED, NKPR_FP32 = 0, 1, 2
NKPR_STOP_STEP, NKPR_STOP_L1_ERROR = 0, 1

# every symbol include/nkpr.h declares
EXPORTED = ["nkpr_params_default", "nkpr_graph_create", "nkpr_graph_create_batch",
            "nkpr_graph_destroy", "nkpr_graph_get_info", "nkpr_graph_stream",
            "nkpr_graph_set_weights", "nkpr_run", "nkpr_run_batch", "nkpr_adjoint",
            "nkpr_ppr_batch", "nkpr_ranking", "nkpr_last_error", "nkpr_abi_version",
            "nkpr_power_params_default", "nkpr_power_run", "nkpr_csr_spmv",
            "nkpr_csr_pagerank_step", "nkpr_algebraic_pagerank", "nkpr_nkbg_read",
            "nkpr_nkbg_free", "nkpr_nkbg_write", "nkpr_csr_pagerank_step_noalloc",
            "nkpr_apr_create", "nkpr_apr_destroy", "nkpr_apr_run", "nkpr_algebraic_create",
            "nkpr_algebraic_run", "nkpr_algebraic_destroy"]
ABI_VERSION = 3


class Csc(ctypes.Structure):
    _fields_ = [("z", ctypes.c_uint64), ("colptr", ctypes.c_void_p), ("src", ctypes.c_void_p),
                ("weight", ctypes.c_void_p), ("weight64", ctypes.c_void_p),
                ("out_wdeg", ctypes.c_void_p), ("exists", ctypes.c_void_p),
                ("device_ptrs", ctypes.c_int32)]


class GraphInfo(ctypes.Structure):
    _fields_ = [("z", ctypes.c_uint64), ("n", ctypes.c_uint64), ("m", ctypes.c_uint64),
                ("sinks", ctypes.c_uint64), ("max_in_degree", ctypes.c_uint64),
                ("work_items", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("weighted", ctypes.c_int32), ("device", ctypes.c_int32),
                ("prep_ms", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("damp", ctypes.c_double), ("tol", ctypes.c_double),
                ("max_iterations", ctypes.c_uint64), ("norm", ctypes.c_int32),
                ("sink_handling", ctypes.c_int32), ("normalized", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("stop", ctypes.c_int32),
                ("scores_on_device", ctypes.c_int32), ("timing", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("interrupt", ctypes.c_void_p),
                ("teleport", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("sweeps_fp32", ctypes.c_uint64),
                ("sweeps_fp64", ctypes.c_uint64), ("refinements", ctypes.c_uint64),
                ("sweeps_fp32_correction", ctypes.c_uint64), ("last_step", ctypes.c_double), ("last_step_l1", ctypes.c_double),
                ("l1_error_bound", ctypes.c_double), ("maximum", ctypes.c_double),
                ("score_sum", ctypes.c_double), ("ms_sweeps_fp32", ctypes.c_double),
                ("ms_sweeps_fp64", ctypes.c_double), ("ms_total", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


NKPR_EIGENVECTOR, NKPR_KATZ = 0, 1


class PowerParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("scores_on_device", ctypes.c_int32),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("tol", ctypes.c_double),
                ("max_iterations", ctypes.c_uint64), ("interrupt", ctypes.c_void_p)]


class CsrMatrix(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_uint64), ("cols", ctypes.c_uint64), ("nnz", ctypes.c_uint64),
                ("row_idx", ctypes.c_void_p), ("col_idx", ctypes.c_void_p),
                ("values", ctypes.c_void_p)]


class Nkbg(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_uint64), ("edges", ctypes.c_uint64),
                ("self_loops", ctypes.c_uint64), ("directed", ctypes.c_int32),
                ("weighted", ctypes.c_int32), ("weight_format", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("colptr", ctypes.c_void_p),
                ("src", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("out_ptr", ctypes.c_void_p), ("out_dst", ctypes.c_void_p),
                ("out_weight", ctypes.c_void_p), ("exists", ctypes.c_void_p)]


class AprStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_uint64), ("pushes", ctypes.c_uint64),
                ("touched", ctypes.c_uint64), ("edge_pushes", ctypes.c_uint64),
                ("ms_total", ctypes.c_double)]


class BatchStats(ctypes.Structure):
    _fields_ = [("sweeps", ctypes.c_uint64), ("sweeps_fp32", ctypes.c_uint64),
                ("sweeps_fp64", ctypes.c_uint64), ("refinements", ctypes.c_uint64),
                ("max_iterations_col", ctypes.c_uint64), ("max_error_bound", ctypes.c_double),
                ("ms_sweeps_fp32", ctypes.c_double), ("ms_sweeps_fp64", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("tile_columns", ctypes.c_uint32),
                ("heavy_degree", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libnkpr.so (raises OSError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libnkpr.so is not built: run __graft_entry__.build()")
    from ._runtime import bind_hip_runtime
    bind_hip_runtime()
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.nkpr_params_default.argtypes = [ctypes.POINTER(Params)]
    L.nkpr_graph_create.argtypes = [ctypes.POINTER(Csc), ctypes.POINTER(vp)]
    L.nkpr_graph_destroy.argtypes = [vp]
    L.nkpr_graph_get_info.argtypes = [vp, ctypes.POINTER(GraphInfo)]
    L.nkpr_graph_stream.argtypes = [vp]
    L.nkpr_graph_stream.restype = vp
    L.nkpr_run.argtypes = [vp, ctypes.POINTER(Params), vp, ctypes.POINTER(Stats)]
    L.nkpr_power_params_default.argtypes = [ctypes.POINTER(PowerParams)]
    L.nkpr_power_run.argtypes = [vp, ctypes.POINTER(PowerParams), vp, ctypes.POINTER(Stats)]
    L.nkpr_graph_create_batch.argtypes = [ctypes.POINTER(Csc), ctypes.c_uint32, vp,
                                          ctypes.POINTER(vp)]
    L.nkpr_run_batch.argtypes = [vp, ctypes.POINTER(Params), vp, vp, vp, ctypes.POINTER(Stats)]
    L.nkpr_graph_set_weights.argtypes = [vp, vp, vp, vp, ctypes.c_int32]
    L.nkpr_adjoint.argtypes = [vp, ctypes.POINTER(Params), ctypes.POINTER(Csc), vp,
                               ctypes.c_double, vp, vp, vp, ctypes.POINTER(Stats)]
    L.nkpr_ppr_batch.argtypes = [vp, ctypes.POINTER(Params), ctypes.c_uint32, vp, vp, vp, vp,
                                 ctypes.POINTER(BatchStats), vp, vp]
    L.nkpr_ranking.argtypes = [ctypes.c_uint64, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_uint64)]
    L.nkpr_csr_spmv.argtypes = [ctypes.POINTER(CsrMatrix), vp, vp, vp]
    L.nkpr_csr_pagerank_step.argtypes = [ctypes.POINTER(CsrMatrix), ctypes.c_double, vp, vp, vp,
                                         vp]
    L.nkpr_csr_pagerank_step_noalloc.argtypes = [ctypes.POINTER(CsrMatrix), ctypes.c_double, vp,
                                                 vp, vp, vp]
    L.nkpr_algebraic_pagerank.argtypes = [ctypes.POINTER(CsrMatrix), ctypes.c_double,
                                          ctypes.c_double, ctypes.c_uint64, vp, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_uint64)]
    L.nkpr_nkbg_read.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(Nkbg)]
    L.nkpr_nkbg_free.argtypes = [ctypes.POINTER(Nkbg)]
    L.nkpr_nkbg_free.restype = None
    L.nkpr_nkbg_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(Nkbg), ctypes.c_uint64,
                                  ctypes.c_int32]
    L.nkpr_apr_create.argtypes = [ctypes.c_uint64, vp, vp, vp, ctypes.c_int32, ctypes.POINTER(vp)]
    L.nkpr_apr_destroy.argtypes = [vp]
    L.nkpr_apr_run.argtypes = [vp, ctypes.c_double, ctypes.c_double, vp, ctypes.c_uint64,
                               ctypes.POINTER(ctypes.c_uint64), vp, vp, vp,
                               ctypes.POINTER(AprStats)]
    L.nkpr_algebraic_create.argtypes = [ctypes.POINTER(CsrMatrix), ctypes.POINTER(vp)]
    L.nkpr_algebraic_run.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, vp,
                                     ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(Stats)]
    L.nkpr_algebraic_destroy.argtypes = [vp]
    L.nkpr_last_error.restype = ctypes.c_char_p
    L.nkpr_abi_version.restype = ctypes.c_int
    for name in EXPORTED:
        if name not in ("nkpr_graph_stream", "nkpr_last_error", "nkpr_nkbg_free"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


class NkprError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nkpr error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != NKPR_OK:
        raise NkprError(rc, lib().nkpr_last_error().decode())


def default_power_params(**kw):
    p = PowerParams()
    check(lib().nkpr_power_params_default(ctypes.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def default_params(**kw):
    p = Params()
    check(lib().nkpr_params_default(ctypes.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def _addr(a):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()  # torch tensor


class DeviceGraph:
    """RAII wrapper of an nkpr_graph handle.

    Host arrays (numpy) are uploaded; device arrays (torch CUDA tensors) are read in place
    (set ``device=True``).  The arrays only have to live until the constructor returns."""

    def __init__(self, z, colptr, src, weight=None, out_wdeg=None, exists=None, weight64=None,
                 device=False, graph_offsets=None):
        self._keep = (colptr, src, weight, out_wdeg, exists, weight64)
        if not device:
            colptr = np.ascontiguousarray(colptr, dtype=np.uint64)
            src = np.ascontiguousarray(src, dtype=np.int32)
            weight = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32)
            weight64 = None if weight64 is None else np.ascontiguousarray(weight64, dtype=np.float64)
            out_wdeg = None if out_wdeg is None else np.ascontiguousarray(out_wdeg, dtype=np.float64)
            exists = None if exists is None else np.ascontiguousarray(exists, dtype=np.uint8)
            self._keep = (colptr, src, weight, out_wdeg, exists, weight64)
        c = Csc(z=int(z), colptr=_addr(colptr), src=_addr(src), weight=_addr(weight),
                weight64=_addr(weight64), out_wdeg=_addr(out_wdeg), exists=_addr(exists),
                device_ptrs=1 if device else 0)
        h = ctypes.c_void_p()
        self.num_graphs = 0
        if graph_offsets is None:
            check(lib().nkpr_graph_create(ctypes.byref(c), ctypes.byref(h)))
        else:
            off = np.ascontiguousarray(graph_offsets, dtype=np.uint64)
            self.num_graphs = len(off) - 1
            check(lib().nkpr_graph_create_batch(ctypes.byref(c), self.num_graphs, _addr(off),
                                                ctypes.byref(h)))
        self.handle = h
        self.z = int(z)
        self.device = device
        self._keep = None

    def set_weights(self, weight, weight64=None, out_wdeg=None):
        """nkpr_graph_set_weights: new weights, same structure (arrays of the constructor's
        residency)."""
        if not self.device:
            weight = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32)
            weight64 = None if weight64 is None else np.ascontiguousarray(weight64, dtype=np.float64)
            out_wdeg = None if out_wdeg is None else np.ascontiguousarray(out_wdeg, dtype=np.float64)
        check(lib().nkpr_graph_set_weights(self.handle, _addr(weight), _addr(weight64),
                                           _addr(out_wdeg), 1 if self.device else 0))

    def adjoint(self, colptr, src, scores, score_sum, grad_scores, weight=None, weight64=None,
                out_wdeg=None, exists=None, want_weight=True, want_teleport=True, params=None,
                **kw):
        """nkpr_adjoint.  (colptr, src, weight...) is the structure the graph was made from,
        with the same residency as the constructor's arrays; scores / grad_scores follow
        params.scores_on_device.  Returns (grad_weight, grad_teleport, stats); host outputs
        are numpy arrays, device outputs must be passed in via kw grad_weight_out /
        grad_teleport_out (torch tensors)."""
        gw_out = kw.pop("grad_weight_out", None)
        gt_out = kw.pop("grad_teleport_out", None)
        p = params if params is not None else default_params(**kw)
        if not self.device:
            colptr = np.ascontiguousarray(colptr, dtype=np.uint64)
            src = np.ascontiguousarray(src, dtype=np.int32)
            weight = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32)
            weight64 = None if weight64 is None else np.ascontiguousarray(weight64, dtype=np.float64)
            out_wdeg = None if out_wdeg is None else np.ascontiguousarray(out_wdeg, dtype=np.float64)
        m = int(colptr[-1])
        c = Csc(z=self.z, colptr=_addr(colptr), src=_addr(src), weight=_addr(weight),
                weight64=_addr(weight64), out_wdeg=_addr(out_wdeg), exists=_addr(exists),
                device_ptrs=1 if self.device else 0)
        gw = gw_out
        if gw is None and want_weight:
            assert not self.device, "pass grad_weight_out for device graphs"
            gw = np.zeros(max(m, 1), dtype=np.float64)
        gt = gt_out
        if gt is None and want_teleport:
            assert not p.scores_on_device, "pass grad_teleport_out for device scores"
            gt = np.zeros(self.z, dtype=np.float64)
        st = Stats()
        check(lib().nkpr_adjoint(self.handle, ctypes.byref(p), ctypes.byref(c), _addr(scores),
                                 float(score_sum), _addr(grad_scores), _addr(gw), _addr(gt),
                                 ctypes.byref(st)))
        if isinstance(gw, np.ndarray):
            gw = gw[:m]
        return gw, gt, st.as_dict()

    def info(self):
        i = GraphInfo()
        check(lib().nkpr_graph_get_info(self.handle, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def stream(self):
        return lib().nkpr_graph_stream(self.handle)

    def run(self, scores, params=None, **kw):
        """nkpr_run on ``scores`` (numpy float64[z] on the host, updated in place, or a torch
        float64 CUDA tensor with params.scores_on_device = 1).  Returns the stats dict."""
        p = params if params is not None else default_params(**kw)
        if isinstance(scores, np.ndarray):
            assert scores.dtype == np.float64 and scores.flags.c_contiguous
            assert len(scores) == self.z
        st = Stats()
        check(lib().nkpr_run(self.handle, ctypes.byref(p), _addr(scores), ctypes.byref(st)))
        return st.as_dict()

    def power_run(self, scores, params=None, **kw):
        """nkpr_power_run (eigenvector or Katz centrality; the handle must have been created
        with out_wdeg = 1).  ``scores`` is numpy float64[z] (output) or a torch CUDA tensor with
        params.scores_on_device = 1.  Returns the stats dict."""
        p = params if params is not None else default_power_params(**kw)
        if isinstance(scores, np.ndarray):
            assert scores.dtype == np.float64 and scores.flags.c_contiguous
            assert len(scores) == self.z
        st = Stats()
        check(lib().nkpr_power_run(self.handle, ctypes.byref(p), _addr(scores), ctypes.byref(st)))
        return st.as_dict()

    def run_batch(self, scores, params=None, **kw):
        """nkpr_run_batch on a batch handle (graph_offsets given at construction).  Returns
        (stats dict, per-graph iterations, per-graph maximum)."""
        p = params if params is not None else default_params(**kw)
        if isinstance(scores, np.ndarray):
            assert scores.dtype == np.float64 and scores.flags.c_contiguous
            assert len(scores) == self.z
        k = max(self.num_graphs, 1)
        its = np.zeros(k, dtype=np.uint64)
        mx = np.zeros(k, dtype=np.float64)
        st = Stats()
        check(lib().nkpr_run_batch(self.handle, ctypes.byref(p), _addr(scores), _addr(its),
                                   _addr(mx), ctypes.byref(st)))
        return st.as_dict(), its[:self.num_graphs], mx[:self.num_graphs]

    def ppr_batch(self, seeds, scores=None, seed_weights=None, params=None, **kw):
        """nkpr_ppr_batch: one personalised PageRank per seed set (a list of k sequences of
        slot ids, optionally with weights).  ``scores`` is the [k][z] start matrix (numpy, or a
        torch CUDA tensor with params.scores_on_device = 1); by default every column starts at
        1/n.  Returns (scores, stats dict, per-column iterations, per-column error bound)."""
        p = params if params is not None else default_params(**kw)
        k = len(seeds)
        off = np.zeros(k + 1, dtype=np.uint64)
        off[1:] = np.cumsum([len(s) for s in seeds])
        nodes = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.int64) for s in seeds])
                                     if k else np.zeros(0), dtype=np.int32)
        wts = None
        if seed_weights is not None:
            wts = np.ascontiguousarray(np.concatenate([np.asarray(w, dtype=np.float64)
                                                       for w in seed_weights]), dtype=np.float64)
            assert len(wts) == len(nodes)
        if scores is None:
            n = self.info()["n"]
            scores = np.full((k, self.z), 1.0 / max(n, 1), dtype=np.float64)
        if isinstance(scores, np.ndarray):
            assert scores.dtype == np.float64 and scores.flags.c_contiguous
            assert scores.shape == (k, self.z)
        st = BatchStats()
        its = np.zeros(max(k, 1), dtype=np.uint64)
        bnd = np.zeros(max(k, 1), dtype=np.float64)
        check(lib().nkpr_ppr_batch(self.handle, ctypes.byref(p), k, _addr(off), _addr(nodes),
                                   _addr(wts), _addr(scores), ctypes.byref(st), _addr(its),
                                   _addr(bnd)))
        return scores, st.as_dict(), its[:k], bnd[:k]

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().nkpr_graph_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ranking(scores, exists=None):
    scores = np.ascontiguousarray(scores, dtype=np.float64)
    z = len(scores)
    ids = np.zeros(max(z, 1), dtype=np.int64)
    vals = np.zeros(max(z, 1), dtype=np.float64)
    k = ctypes.c_uint64(0)
    ex = None if exists is None else np.ascontiguousarray(exists, dtype=np.uint8)
    check(lib().nkpr_ranking(z, _addr(scores), _addr(ex), _addr(ids), _addr(vals),
                             ctypes.byref(k)))
    return ids[:k.value], vals[:k.value]


class DeviceCSR:
    """A CSRGeneralMatrix<double> in device memory (torch CUDA tensors: row_idx and col_idx
    int64 viewed as NetworKit's uint64 index, values float64), for nkpr_csr_spmv,
    nkpr_csr_pagerank_step and nkpr_algebraic_pagerank.  ``row_idx`` etc. may be numpy
    arrays; they are copied to the current device."""

    def __init__(self, rows, cols, row_idx, col_idx, values):
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())

        def on_dev(a, dt):
            if isinstance(a, np.ndarray):
                a = torch.from_numpy(np.ascontiguousarray(a).view(np.int64) if a.dtype == np.uint64
                                     else np.ascontiguousarray(a))
            return a.to(device=dev, dtype=dt).contiguous()

        self.row_idx = on_dev(row_idx, torch.int64)
        self.col_idx = on_dev(col_idx, torch.int64)
        self.values = on_dev(values, torch.float64)
        self.rows, self.cols = int(rows), int(cols)
        self.nnz = int(self.col_idx.numel())
        assert self.row_idx.numel() == self.rows + 1 and self.values.numel() == self.nnz

    def struct(self):
        return CsrMatrix(rows=self.rows, cols=self.cols, nnz=self.nnz,
                         row_idx=self.row_idx.data_ptr(), col_idx=self.col_idx.data_ptr(),
                         values=self.values.data_ptr())

    def spmv(self, x, y=None, stream=None):
        """y = A x on the device (x a float64 CUDA tensor of cols entries)."""
        import torch
        if y is None:
            y = torch.empty(self.rows, dtype=torch.float64, device=x.device)
        m = self.struct()
        check(lib().nkpr_csr_spmv(ctypes.byref(m), x.data_ptr(), y.data_ptr(), stream))
        return y

    def pagerank_step(self, x, shift, y=None, step_sq=None, stream=None):
        """y = M x + shift and ||y - x||^2 (a one-element float64 CUDA tensor)."""
        import torch
        if y is None:
            y = torch.empty(self.rows, dtype=torch.float64, device=x.device)
        if step_sq is None:
            step_sq = torch.empty(1, dtype=torch.float64, device=x.device)
        m = self.struct()
        check(lib().nkpr_csr_pagerank_step(ctypes.byref(m), float(shift), x.data_ptr(),
                                           y.data_ptr(), step_sq.data_ptr(), stream))
        return y, step_sq

    def pagerank_step_noalloc(self, x, shift, y, step_sq, stream=None):
        """nkpr_csr_pagerank_step_noalloc: the step into caller-provided y and step_sq (no
        allocation by the library)."""
        m = self.struct()
        check(lib().nkpr_csr_pagerank_step_noalloc(ctypes.byref(m), float(shift), x.data_ptr(),
                                                   y.data_ptr(), step_sq.data_ptr(), stream))
        return y, step_sq

    def algebraic_pagerank(self, damp=0.85, tol=1e-8, max_iterations=2**64 - 1, scores=None):
        """nkpr_algebraic_pagerank on this matrix as M; returns (scores, iterations) with
        scores a numpy array (or the given CUDA tensor, filled in place)."""
        m = self.struct()
        it = ctypes.c_uint64(0)
        if scores is None:
            out = np.zeros(self.rows, dtype=np.float64)
            check(lib().nkpr_algebraic_pagerank(ctypes.byref(m), damp, tol, int(max_iterations),
                                                _addr(out), 0, ctypes.byref(it)))
            return out, int(it.value)
        check(lib().nkpr_algebraic_pagerank(ctypes.byref(m), damp, tol, int(max_iterations),
                                            scores.data_ptr(), 1, ctypes.byref(it)))
        return scores, int(it.value)


class DeviceAlgebraic:
    """nkpr_algebraic handle over a DeviceCSR's matrix M (built once, reused by every run)."""

    def __init__(self, csr):
        m = csr.struct()
        h = ctypes.c_void_p()
        check(lib().nkpr_algebraic_create(ctypes.byref(m), ctypes.byref(h)))
        self.handle = h
        self.rows = csr.rows

    def run(self, damp=0.85, tol=1e-8, max_iterations=2**64 - 1, scores=None, stats=False):
        """nkpr_algebraic_run; returns (scores, iterations) like DeviceCSR.algebraic_pagerank,
        and the stats dict third when stats=True (sweeps timed with HIP events)."""
        it = ctypes.c_uint64(0)
        st = Stats()
        sp = ctypes.byref(st) if stats else None
        if scores is None:
            out = np.zeros(self.rows, dtype=np.float64)
            check(lib().nkpr_algebraic_run(self.handle, damp, tol, int(max_iterations), _addr(out),
                                           0, ctypes.byref(it), sp))
        else:
            out = scores
            check(lib().nkpr_algebraic_run(self.handle, damp, tol, int(max_iterations),
                                           scores.data_ptr(), 1, ctypes.byref(it), sp))
        if stats:
            return out, int(it.value), {f: getattr(st, f) for f, _ in st._fields_}
        return out, int(it.value)

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().nkpr_algebraic_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _copy(ptr, dtype, count):
    if not ptr or count == 0:
        return np.zeros(count, dtype=dtype)
    ct = {np.uint64: ctypes.c_uint64, np.int32: ctypes.c_int32, np.float64: ctypes.c_double,
          np.uint8: ctypes.c_uint8}[dtype]
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(count,)).copy()


def read_nkbg(path, with_out_lists=False):
    """nkpr_nkbg_read: a NetworKit binary graph file as flat host lists (numpy copies).
    Returns a dict: nodes, edges, self_loops, directed, weighted, weight_format, colptr, src,
    weight (None when unweighted), exists, and out_ptr / out_dst / out_weight for a directed
    graph when with_out_lists."""
    g = Nkbg()
    check(lib().nkpr_nkbg_read(os.fsencode(path), 1 if with_out_lists else 0, ctypes.byref(g)))
    try:
        n = g.nodes
        m = _copy(g.colptr, np.uint64, n + 1)[-1] if n else 0
        out = {"nodes": n, "edges": g.edges, "self_loops": g.self_loops,
               "directed": bool(g.directed), "weighted": bool(g.weighted),
               "weight_format": g.weight_format,
               "colptr": _copy(g.colptr, np.uint64, n + 1),
               "src": _copy(g.src, np.int32, int(m)),
               "weight": _copy(g.weight, np.float64, int(m)) if g.weighted else None,
               "exists": _copy(g.exists, np.uint8, n)}
        if g.out_ptr:
            op = _copy(g.out_ptr, np.uint64, n + 1)
            out["out_ptr"] = op
            out["out_dst"] = _copy(g.out_dst, np.int32, int(op[-1]))
            out["out_weight"] = _copy(g.out_weight, np.float64, int(op[-1])) if g.weighted else None
        return out
    finally:
        lib().nkpr_nkbg_free(ctypes.byref(g))


def write_nkbg(path, nodes, directed, out_ptr, out_dst, out_weight=None, in_ptr=None,
               in_src=None, in_weight=None, exists=None, chunks=32, weight_format=-1):
    """nkpr_nkbg_write from host lists in storage order (see nkpr.h)."""
    keep = []

    def arr(a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    g = Nkbg(nodes=int(nodes), directed=1 if directed else 0,
             weighted=0 if out_weight is None else 1,
             out_ptr=arr(out_ptr, np.uint64), out_dst=arr(out_dst, np.int32),
             out_weight=arr(out_weight, np.float64), colptr=arr(in_ptr, np.uint64),
             src=arr(in_src, np.int32), weight=arr(in_weight, np.float64),
             exists=arr(exists, np.uint8))
    check(lib().nkpr_nkbg_write(os.fsencode(path), ctypes.byref(g), int(chunks),
                                int(weight_format)))


class DeviceAPR:
    """nkpr_apr handle over out-lists (host numpy arrays or, with device=True, torch CUDA
    tensors): ptr uint64[z+1], dst int32[m], w float64[m] or None."""

    def __init__(self, z, ptr, dst, w=None, device=False):
        if not device:
            ptr = np.ascontiguousarray(ptr, dtype=np.uint64)
            dst = np.ascontiguousarray(dst, dtype=np.int32)
            w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().nkpr_apr_create(int(z), _addr(ptr), _addr(dst), _addr(w),
                                    1 if device else 0, ctypes.byref(h)))
        self.handle = h
        self.z = int(z)

    def run(self, seeds, alpha, eps=1e-12):
        """Returns (ids, p, r, stats): the touched nodes ascending, their scores and residuals."""
        sd = np.ascontiguousarray(np.atleast_1d(np.asarray(seeds)), dtype=np.uint64)
        ids = np.zeros(self.z, dtype=np.int64)
        p = np.zeros(self.z, dtype=np.float64)
        r = np.zeros(self.z, dtype=np.float64)
        k = ctypes.c_uint64(0)
        st = AprStats()
        check(lib().nkpr_apr_run(self.handle, float(alpha), float(eps), _addr(sd), len(sd),
                                 ctypes.byref(k), _addr(ids), _addr(p), _addr(r),
                                 ctypes.byref(st)))
        n = k.value
        return ids[:n], p[:n], r[:n], {f: getattr(st, f) for f, _ in st._fields_}

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().nkpr_apr_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
