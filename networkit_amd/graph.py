"""networkit.graph equivalent: the Graph class (networkit/graph.pyx, subset)."""
from ._core import Graph  # noqa: F401

__all__ = ["Graph"]
