"""networkit.base equivalent: the Algorithm base class (networkit/base.pyx)."""
from ._core import Algorithm  # noqa: F401

__all__ = ["Algorithm"]
