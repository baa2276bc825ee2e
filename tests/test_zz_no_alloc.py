"""fsx_cg_solve_noalloc (no counterpart in the reference) against fsx_cg_solve
on the same inputs (gpu).

The no-allocation entry point keeps the CG state, setup scalars, reduction
partials and work vectors in the library's static per-device scratch array
and launches the iterations directly; fsx_cg_solve uses pooled context
buffers and captured graphs.  Both run the same kernels in the same order
over the same data, so the solution, the iteration count and every
solve_info field must agree bit for bit.  The calls run back to back, each
later one on a smaller system than the one before (the array keeps the last
call's contents), and none may make a device allocation.
"""
import numpy as np
import pytest

from flecsolve_amd import fsx
from flecsolve_amd import problems as P

pytestmark = pytest.mark.gpu


def diffusion(ctx, n, seed):
    k = P.lognormal_field(n, 1.0, 8.0, seed)
    rp, ci, v = P.diffusion3d(n, k)
    return ctx.csr(rp, ci, v)


def same_info(a, b):
    assert a.as_dict() == b.as_dict(), (a.as_dict(), b.as_dict())


def run_both(ctx, A, b, x0, s, dinv=None, dtype=fsx.F64):
    """x from fsx_cg_solve and from fsx_cg_solve_noalloc, same x0; the
    allocations the no-allocation call made."""
    xa = ctx.vec(x0, dtype=dtype)
    ia = fsx.cg_solve(A, ctx.vec(b, dtype=dtype), xa, s, dinv=dinv)
    bn = ctx.vec(b, dtype=dtype)
    xn = ctx.vec(x0, dtype=dtype)
    before = fsx.alloc_stats()[0]
    inn = fsx.cg_solve_noalloc(A, bn, xn, s, dinv=dinv)
    made = fsx.alloc_stats()[0] - before
    return xa.get(), ia, xn.get(), inn, made


def test_noalloc_matches_cg_solve_two_calls(ctx):
    # call 1: 96^3 = 884,736 rows, Jacobi, zero initial guess, FP64
    n1 = 96
    A1 = diffusion(ctx, n1, 2001)
    d1 = A1.diag_inv()
    b1 = P.uniform(n1 ** 3, 2002)
    s1 = fsx.Settings(maxiter=4000, rtol=1e-10, use_zero_guess=True)
    xa, ia, xn, inn, _ = run_both(ctx, A1, b1, np.zeros(n1 ** 3), s1, dinv=d1)
    print(f"noalloc 96^3 jacobi: iters {inn.iters} ({inn.reason})", flush=True)
    assert ia.reason == "converged_rtol"
    same_info(ia, inn)
    assert np.array_equal(xa, xn)

    # call 2, smaller: 40^3, identity preconditioner, nonzero initial guess
    # (the setup's residual product b - A x0 runs in the scratch array too)
    n2 = 40
    A2 = diffusion(ctx, n2, 2003)
    b2 = P.uniform(n2 ** 3, 2004)
    x0 = P.uniform(n2 ** 3, 2005)
    s2 = fsx.Settings(maxiter=3000, rtol=1e-8, use_zero_guess=False)
    xa, ia, xn, inn, made = run_both(ctx, A2, b2, x0, s2)
    print(f"noalloc 40^3 identity: iters {inn.iters} ({inn.reason})", flush=True)
    assert ia.reason == "converged_rtol"
    same_info(ia, inn)
    assert np.array_equal(xa, xn)
    assert made == 0, f"{made} device allocations"

    # call 3, smaller still: FP32 matrix and vectors, Jacobi (FP32 D^-1)
    n3 = 24
    A3_64 = diffusion(ctx, n3, 2006)
    A3 = A3_64.convert(fsx.F32)
    d3 = A3_64.diag_inv(fsx.F32)
    b3 = P.uniform(n3 ** 3, 2007).astype(np.float32)
    s3 = fsx.Settings(maxiter=2000, rtol=1e-5, use_zero_guess=True)
    xa, ia, xn, inn, made = run_both(ctx, A3, b3, np.zeros(n3 ** 3, np.float32), s3,
                                     dinv=d3, dtype=fsx.F32)
    print(f"noalloc 24^3 fp32: iters {inn.iters} ({inn.reason})", flush=True)
    same_info(ia, inn)
    assert np.array_equal(xa, xn)
    assert made == 0, f"{made} device allocations"

    # maxiter reached (iters reported 0, diverged_iters) and the early exit
    # of a zero right-hand side, both on the 40^3 system
    for s, b in [(fsx.Settings(maxiter=7, rtol=1e-12, use_zero_guess=True), b2),
                 (fsx.Settings(maxiter=50, rtol=1e-8, use_zero_guess=True),
                  np.zeros(n2 ** 3))]:
        xa, ia, xn, inn, made = run_both(ctx, A2, b, np.zeros(n2 ** 3), s)
        same_info(ia, inn)
        assert np.array_equal(xa, xn)
        assert made == 0


def test_noalloc_rejects_oversize(ctx):
    n = fsx.NOALLOC_MAX_ROWS + 1
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    A = ctx.csr(rp, ci, np.full(n, 2.0))
    b = ctx.vec(np.ones(n))
    x = ctx.zeros(n)
    with pytest.raises(fsx.FsxError, match="scratch limit"):
        fsx.cg_solve_noalloc(A, b, x, fsx.Settings(maxiter=10, rtol=1e-8,
                                                   use_zero_guess=True))
    # the same system at the limit's size is accepted
    m = fsx.NOALLOC_MAX_ROWS
    A = ctx.csr(rp[:m + 1], ci[:m], np.full(m, 2.0))
    x = ctx.zeros(m)
    info = fsx.cg_solve_noalloc(A, ctx.vec(np.ones(m)), x,
                                fsx.Settings(maxiter=10, rtol=1e-8, use_zero_guess=True))
    assert info.reason == "converged_rtol" and info.iters == 1
    assert np.array_equal(x.get(), np.full(m, 0.5))
