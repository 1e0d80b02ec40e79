"""networkit_amd -- an MI355X-native drop-in for NetworKit's PageRank.

    import networkit_amd as nk
    G = nk.Graph(n, weighted, directed); G.addEdge(u, v, w)
    pr = nk.centrality.PageRank(G, damp=0.85, tol=1e-8)
    pr.run(); pr.scores(); pr.ranking()

The compute path is native only: libnkpr.so (hand-written HIP for gfx950 behind the C ABI of
include/nkpr.h) and the Cython extension _core that wraps the C++ drop-in classes.  Importing
this package without those built fails with ImportError; there is no CPU fallback.
"""
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

from ._runtime import bind_hip_runtime as _bind_hip_runtime  # noqa: E402

_bind_hip_runtime()

try:
    from . import _core  # noqa: F401
except ImportError as _e:  # pragma: no cover - exercised only when the build is missing
    raise ImportError(
        "networkit_amd's native extension is not built (run __graft_entry__.build() or "
        "python -m networkit_amd._build): %s" % _e) from _e

from . import algebraic, base, centrality, graph, graphio, scd  # noqa: E402,F401
from .graph import Graph  # noqa: E402,F401

__all__ = ["Graph", "algebraic", "base", "centrality", "graph", "graphio", "scd"]
