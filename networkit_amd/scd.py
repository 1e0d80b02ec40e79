"""Selective community detection on the PageRank path (networkit.scd equivalent, scd.pyx:93-156).

ApproximatePageRank and PageRankNibble run their residual pushes on the MI355X
(``nkpr_apr_run``); PageRankNibble's sweep over the result runs on the host as in the
reference (networkit/cpp/scd/PageRankNibble.cpp:28-89).
"""
from ._core import ApproximatePageRank, PageRankNibble  # noqa: F401

__all__ = ["ApproximatePageRank", "PageRankNibble"]
