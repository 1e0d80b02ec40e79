"""The algebraic module's PageRank on the MI355X.

AlgebraicPageRank is NetworKit's C++ ``AlgebraicPageRank<CSRMatrix>``
(include/networkit/algebraic/algorithms/AlgebraicPageRank.hpp:25-150) through the C++ drop-in
in ``csrc/AlgebraicPageRank.{hpp,cpp}``; the CSR product behind it is
``CSRGeneralMatrix::operator*(const Vector&)`` (CSRGeneralMatrix.hpp:563-578), exported as
``nkpr_csr_spmv`` (``capi.DeviceCSR`` binds it for device arrays).
"""
from ._core import AlgebraicPageRank  # noqa: F401

__all__ = ["AlgebraicPageRank"]
