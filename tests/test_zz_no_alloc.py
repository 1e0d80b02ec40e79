"""nkpr_csr_pagerank_step_noalloc (the allocation-free variant of nkpr_csr_pagerank_step,
whose partial sums live in a device array of the library) against the allocating entry
point, through the C ABI: the same y and ||y - x||^2 bit for bit (both add their partials in
an order fixed by the matrix), on two calls in a row, the second on a smaller matrix, so that
the scratch's contents from the first call cannot leak into the second."""
import numpy as np
import pytest

import oracle
from networkit_amd import capi

pytestmark = pytest.mark.gpu


def matrix(rows, avg, long_rows, seed):
    rng = np.random.default_rng(seed)
    lengths = rng.poisson(avg, rows)
    if long_rows:
        lengths[rng.choice(rows, long_rows, replace=False)] = rng.integers(2049, 9000, long_rows)
    row_idx = np.zeros(rows + 1, np.uint64)
    row_idx[1:] = np.cumsum(lengths)
    nnz = int(row_idx[-1])
    col = rng.integers(0, rows, nnz).astype(np.uint64)
    return oracle.CSR(rows, rows, row_idx, col, rng.random(nnz) / max(avg, 1))


def compare(a, seed):
    import torch
    d = capi.DeviceCSR(a.rows, a.cols, a.row_idx, a.col_idx, a.values)
    x = torch.from_numpy(np.random.default_rng(seed).random(a.rows)).cuda()
    y_ref, sq_ref = d.pagerank_step(x, 0.15 / a.rows)
    y = torch.full_like(y_ref, float("nan"))
    sq = torch.full((1,), float("nan"), dtype=torch.float64, device=x.device)
    d.pagerank_step_noalloc(x, 0.15 / a.rows, y, sq)
    torch.cuda.synchronize()
    assert torch.equal(y, y_ref)
    assert float(sq.item()) == float(sq_ref.item())
    # and both against the oracle's product
    ref = oracle.csr_spmv(a, x.cpu().numpy()) + 0.15 / a.rows
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-12, atol=1e-15)
    return float(sq.item())


def test_noalloc_step_matches_allocating_step_twice():
    big = matrix(200_000, 14, 40, 1)      # both kernels, many blocks: most of the scratch
    small = matrix(3_000, 5, 2, 2)        # fewer blocks: a prefix of the scratch
    s1 = compare(big, 3)
    s2 = compare(small, 4)
    assert s1 > 0 and s2 > 0
    # a third call on the first matrix again
    assert compare(big, 3) == s1


def test_noalloc_step_rejects_bad_arguments():
    import torch
    a = matrix(100, 4, 0, 5)
    d = capi.DeviceCSR(a.rows, a.cols, a.row_idx, a.col_idx, a.values)
    x = torch.ones(a.rows, dtype=torch.float64, device="cuda")
    sq = torch.empty(1, dtype=torch.float64, device="cuda")
    with pytest.raises(capi.NkprError, match="different"):
        d.pagerank_step_noalloc(x, 0.0, x, sq)


def test_noalloc_step_rejects_matrices_beyond_its_scratch():
    """A matrix that could need more long-row pieces than the scratch holds is refused before
    any kernel runs (the struct claims 2^35 entries, more than 2^22 pieces; no array is read)."""
    import ctypes
    import torch
    a = matrix(100, 4, 0, 6)
    d = capi.DeviceCSR(a.rows, a.cols, a.row_idx, a.col_idx, a.values)
    m = d.struct()
    m.nnz = 1 << 35
    x = torch.ones(a.rows, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    sq = torch.empty(1, dtype=torch.float64, device="cuda")
    rc = capi.lib().nkpr_csr_pagerank_step_noalloc(ctypes.byref(m), 0.0, x.data_ptr(),
                                                   y.data_ptr(), sq.data_ptr(), None)
    assert rc == capi.NKPR_ERR_INVALID
    assert b"too large" in capi.lib().nkpr_last_error()
