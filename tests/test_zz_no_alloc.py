"""bem_lu_solve_noalloc, the allocation-free variant of bem_lu_solve, through
the C ABI: bitwise the same solution as bem_lu_solve on two calls in a row
(the second smaller and of another dtype), the size limit of its fixed
scratch, and capture in a HIP graph (a capture fails on any allocation or
synchronisation inside it), replayed with eager work between the replays."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _factored(bem, ctx, n, dtype, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    if dtype.is_complex:
        a = a + 1j * rng.standard_normal((n, n))
    A = torch.tensor(a, dtype=dtype, device=DEV)
    LU = A.clone()
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    ctx.lu_factor(LU, ipiv)
    b = torch.tensor(rng.standard_normal(n) + (1j * rng.standard_normal(n) if dtype.is_complex else 0),
                     dtype=dtype, device=DEV)
    return A, LU, ipiv, b


def _noalloc(bem, ctx, LU, ipiv, b):
    """The raw C call, as a caller without the Python wrapper makes it."""
    L = bem._lib
    rc = L.lib().bem_lu_solve_noalloc(ctx.handle, LU.data_ptr(), int(LU.stride(0)), int(LU.shape[0]),
                                      L.dtype_code(LU), ipiv.data_ptr(), b.data_ptr(), L.current_stream())
    L.check(rc)


def test_noalloc_solve_matches_lu_solve(bem):
    ctx = bem.Context(0)
    for n, dtype, seed in ((2100, torch.complex128, 1), (300, torch.complex64, 2), (130, torch.float64, 3)):
        print(f"[noalloc] n={n} dtype={dtype}", flush=True)
        A, LU, ipiv, b = _factored(bem, ctx, n, dtype, seed)
        x_ref = b.clone()
        ctx.lu_solve(LU, ipiv, x_ref)
        x = b.clone()
        _noalloc(bem, ctx, LU, ipiv, x)
        torch.cuda.synchronize()
        assert torch.equal(x, x_ref), n
        r = (torch.linalg.vector_norm(A.to(torch.complex128) @ x.to(torch.complex128) - b.to(torch.complex128)) /
             torch.linalg.vector_norm(b.to(torch.complex128))).item()
        assert r <= (1e-12 if dtype in (torch.complex128, torch.float64) else 1e-4)
    ctx.close()


def test_noalloc_solve_rejects_sizes_beyond_its_scratch(bem):
    """n above the scratch limit is refused before any device access (the
    pointers below are valid but far too small for the stated n)."""
    L = bem._lib
    ctx = bem.Context(0)
    buf = torch.zeros(4, dtype=torch.complex128, device=DEV)
    ipiv = torch.zeros(4, dtype=torch.int32, device=DEV)
    n = 65537
    rc = L.lib().bem_lu_solve_noalloc(ctx.handle, buf.data_ptr(), n, n, L.C128, ipiv.data_ptr(), buf.data_ptr(),
                                      L.current_stream())
    assert rc == L.BEM_ERR_INVALID
    with pytest.raises(bem.BEMError, match="65536"):
        L.check(rc)
    ctx.close()


def test_noalloc_solve_in_a_graph(bem):
    ctx = bem.Context(0)
    n = 700
    A, LU, ipiv, b0 = _factored(bem, ctx, n, torch.complex128, 4)
    x_ref = b0.clone()
    ctx.lu_solve(LU, ipiv, x_ref)
    x = b0.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        _noalloc(bem, ctx, LU, ipiv, x)  # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    x.copy_(b0)
    with torch.cuda.graph(g):
        _noalloc(bem, ctx, LU, ipiv, x)
    for rep in range(3):
        x.copy_(b0)  # eager work between the replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(x, x_ref), rep
    del g
    ctx.close()
