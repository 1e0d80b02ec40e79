"""Differentiable PageRank: a torch.autograd.Function over the device library.

Forward: the sum-normalised PageRank scores x(w, t) of a fixed graph structure with edge
weights w (one per in-edge, laid out like the structure's CSC) and a teleport distribution t
(None: the reference's uniform (1-d)/n), as NetworKit's PageRank::run computes them with
normalized = false and no sink handling (networkit/cpp/centrality/PageRank.cpp:20-92).
Backward: nkpr_adjoint, which solves lambda = d P^T lambda + g on the transposed operator and
returns dL/dw and dL/dt (no counterpart in the reference).

There is no CPU path: the HIP library must be built and a gfx950 device present.
"""
import torch

from . import capi


class PageRankOperator:
    """A graph structure on the device: in-lists (colptr[z+1], src[m]) in the layout of
    nkpr_csc.  Call it with edge weights (and optionally a teleport vector) to get scores that
    autograd can differentiate.

    precision: capi.NKPR_MIXED (default; certified L1 error <= tol), NKPR_FP64, or NKPR_FP32.
    The adjoint runs at the same precision and tolerance (relative residual)."""

    def __init__(self, colptr, src, damp=0.85, tol=1e-8, precision=capi.NKPR_MIXED,
                 adjoint_tol=None, device="cuda"):
        self.device = torch.device(device)
        self.colptr = torch.as_tensor(colptr, dtype=torch.int64).to(self.device).contiguous()
        self.src = torch.as_tensor(src, dtype=torch.int32).to(self.device).contiguous()
        self.z = int(self.colptr.numel() - 1)
        self.m = int(self.colptr[-1].item())
        self.damp = float(damp)
        self.tol = float(tol)
        self.adjoint_tol = float(tol if adjoint_tol is None else adjoint_tol)
        self.precision = int(precision)
        ones = torch.ones(max(self.m, 1), dtype=torch.float32, device=self.device)
        torch.cuda.synchronize(self.device)
        self.graph = capi.DeviceGraph(self.z, self.colptr, self.src, ones, device=True)
        self.max_iterations = 100000
        self.last_forward = None
        self.last_backward = None
        # the device handle holds one set of weights: the version of the forward call that
        # loaded them, so that a backward of an earlier call can load its own weights again
        self._calls = 0
        self._loaded = -1

    def _set_weights(self, weight):
        if weight.numel() != self.m:
            raise ValueError("expected %d edge weights, got %d" % (self.m, weight.numel()))
        w = weight.detach()
        if w.dtype == torch.float64:
            self._w64 = w.contiguous()
            self._w32 = self._w64.float()
        else:
            self._w32 = w.float().contiguous()
            self._w64 = None
        torch.cuda.current_stream(self.device).synchronize()
        self.graph.set_weights(self._w32, self._w64)
        self._calls += 1
        self._loaded = self._calls
        return self._w32, self._w64, self._loaded

    def _ensure_weights(self, w32, w64, version):
        """Load the weights of forward call `version` again when a later call replaced them:
        the adjoint's transposed operator is refolded from the handle's current weights."""
        if self._loaded == version:
            return
        torch.cuda.current_stream(self.device).synchronize()
        self.graph.set_weights(w32, w64)
        self._loaded = version

    def params(self, tol, **kw):
        stop = capi.NKPR_STOP_STEP if self.precision == capi.NKPR_FP64 else capi.NKPR_STOP_L1_ERROR
        return capi.default_params(damp=self.damp, tol=tol, precision=self.precision,
                                   stop=stop, norm=capi.NKPR_L1, scores_on_device=1,
                                   max_iterations=self.max_iterations, **kw)

    def __call__(self, weight, teleport=None):
        return PageRankFunction.apply(weight, teleport, self)


class PageRankFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weight, teleport, op):
        w32, w64, version = op._set_weights(weight)
        x = torch.full((op.z,), 1.0 / op.z, dtype=torch.float64, device=op.device)
        t = None
        if teleport is not None:
            t = teleport.detach().to(device=op.device, dtype=torch.float64).contiguous()
        torch.cuda.current_stream(op.device).synchronize()
        p = op.params(op.tol)
        if t is not None:
            p.teleport = t.data_ptr()
        st = op.graph.run(x, p)
        op.last_forward = st
        ctx.op = op
        ctx.score_sum = st["score_sum"]
        ctx.has_teleport = teleport is not None
        ctx.w = (w32, w64)
        ctx.version = version
        ctx.save_for_backward(x)
        return x.to(weight.dtype if weight.dtype.is_floating_point else torch.float64)

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        op = ctx.op
        w32, w64 = ctx.w
        op._ensure_weights(w32, w64, ctx.version)
        c = grad_x.detach().to(torch.float64).contiguous()
        gw = torch.empty(max(op.m, 1), dtype=torch.float64, device=op.device)
        gt = torch.empty(op.z, dtype=torch.float64, device=op.device)
        torch.cuda.current_stream(op.device).synchronize()
        _, _, st = op.graph.adjoint(op.colptr, op.src, x, ctx.score_sum, c, weight=w32,
                                    weight64=w64, grad_weight_out=gw, grad_teleport_out=gt,
                                    params=op.params(op.adjoint_tol))
        op.last_backward = st
        need_w, need_t = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        grad_w = gw[:op.m].to(w64.dtype if w64 is not None else w32.dtype) if need_w else None
        grad_t = gt if (need_t and ctx.has_teleport) else None
        return grad_w, grad_t, None


def pagerank(weight, colptr=None, src=None, teleport=None, op=None, **kw):
    """Functional form: scores = pagerank(weight, colptr, src[, teleport]) (builds a
    PageRankOperator when `op` is not given)."""
    if op is None:
        op = PageRankOperator(colptr, src, **kw)
    return op(weight, teleport)
