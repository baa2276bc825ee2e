"""fsx_ir_cg_rr_solve, the FP32 CG with FP64 residual replacement, against
its CPU model (tests/rr_model.py) and the oracle (gpu).

  * 32^3 lognormal diffusion at sigma 1 and 2.5: true FP64 residual
    <= 1e-10, the model's segment count (+-1) and FP32 iteration count
    (+-3; device and model round the FP32 products and the FP64 sums in
    other orders), solution within the conditioning bound of the oracle's
    FP64 CG + Dinv, and fewer iterations than the restarted refinement
    (fsx_ir_cg_solve, jacobi = 2) on the same FP32 matrix;
  * max_segments = 1: diverged_iters after one segment, x within FP32
    rounding of the model's x after its one segment;
  * a nonzero initial guess; a guess that already meets the tolerance (no
    segment); two calls in a row give the same x bit for bit, also after a
    restarted refinement on the same context (shared work vectors and
    captured graphs);
  * argument errors.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import rr_model as M
from flecsolve_amd import fsx
from flecsolve_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def problem(ctx, n, sigma, seed=3001):
    k = P.lognormal_field(n, sigma, 8.0, seed)
    rp, ci, v = P.diffusion3d(n, k)
    b = P.uniform(n ** 3, seed + 1)
    A64 = ctx.csr(rp, ci, v)
    A32 = A64.convert(fsx.F32)
    A32.set_values_sym_scaled(A64)
    return rp, ci, v, b, A64, A32


def true_res(rp, ci, v, b, x):
    return np.linalg.norm(b - sp.csr_matrix((v, ci, rp)) @ x) / np.linalg.norm(b)


@pytest.mark.parametrize("sigma", [1.0, 2.5])
def test_rr_against_model_and_oracle(ctx, sigma):
    n = 32
    rp, ci, v, b, A64, A32 = problem(ctx, n, sigma)
    xd = ctx.zeros(n ** 3)
    info = fsx.ir_cg_rr_solve(A64, A32, ctx.vec(b), xd, 1e-10,
                              fsx.Settings(maxiter=2000, rtol=1e-2))
    _, msegs, mits, mst, _ = M.rr_solve(rp, ci, v, b, 1e-10, 1e-2, 2000, 100)
    print("device", info.outer_iters, info.inner_iters_total, "model", msegs, mits)
    assert info.status == mst == 1
    assert info.rel_res_final <= 1e-10
    x = xd.get()
    assert true_res(rp, ci, v, b, x) <= 1.01e-10
    assert abs(info.outer_iters - msegs) <= 1
    assert abs(info.inner_iters_total - mits) <= 3
    Ao = O.Csr(rp, ci, v)
    xo, io = O.cg(Ao, b, np.zeros_like(b), O.Settings(maxiter=5000, rtol=1e-12,
                                                     use_zero_guess=True),
                  P=O.prec(1, Ao.dinv()))[:2]
    assert rel(x, xo) < 1e-6  # both 1e-10-residual solutions: cond * 1e-10
    xr = ctx.zeros(n ** 3)
    ir = fsx.ir_cg_solve(A64, A32, ctx.vec(b), xr, 1e-10,
                         fsx.Settings(maxiter=2000, rtol=1e-4), jacobi=2)
    assert ir.status == 1
    assert info.inner_iters_total < ir.inner_iters_total


def test_rr_one_segment_matches_model(ctx):
    n = 24
    rp, ci, v, b, A64, A32 = problem(ctx, n, 1.0, seed=3101)
    xd = ctx.zeros(n ** 3)
    info = fsx.ir_cg_rr_solve(A64, A32, ctx.vec(b), xd, 1e-10,
                              fsx.Settings(maxiter=2000, rtol=1e-2), max_segments=1)
    xm, msegs, mits, mst, _ = M.rr_solve(rp, ci, v, b, 1e-10, 1e-2, 2000, 1)
    assert (info.status, info.outer_iters) == (mst, msegs) == (4, 1)
    assert abs(info.inner_iters_total - mits) <= 2
    assert info.rel_res_final > 1e-10
    # FP32 rounding apart, the same x; a count one or two apart moves x by
    # the last updates (a wrong scale of the x update would be off by O(1))
    assert rel(xd.get(), xm) < (1e-3 if info.inner_iters_total == mits else 5e-2)


def test_rr_guesses_and_repeat(ctx):
    n = 24
    rp, ci, v, b, A64, A32 = problem(ctx, n, 2.0, seed=3201)
    N = n ** 3
    bd = ctx.vec(b)
    s = fsx.Settings(maxiter=2000, rtol=1e-2)
    # a nonzero guess
    x0 = P.uniform(N, 3203)
    xd = ctx.vec(x0)
    info = fsx.ir_cg_rr_solve(A64, A32, bd, xd, 1e-10, s)
    assert info.status == 1 and true_res(rp, ci, v, b, xd.get()) <= 1.01e-10
    _, msegs, mits, _, _ = M.rr_solve(rp, ci, v, b, 1e-10, 1e-2, 2000, 100, x0=x0)
    assert abs(info.inner_iters_total - mits) <= 3
    # a guess that meets the tolerance: no segment
    again = fsx.ir_cg_rr_solve(A64, A32, bd, xd, 1e-9, s)
    assert (again.status, again.outer_iters, again.inner_iters_total) == (1, 0, 0)
    # repeatable, also after a restarted refinement on the same context
    outs = []
    for k in range(3):
        x = ctx.zeros(N)
        if k == 1:
            fsx.ir_cg_solve(A64, A32, bd, ctx.zeros(N), 1e-10,
                            fsx.Settings(maxiter=2000, rtol=1e-4), jacobi=2)
        i = fsx.ir_cg_rr_solve(A64, A32, bd, x, 1e-10, s)
        outs.append((i.outer_iters, i.inner_iters_total, x.get()))
    for segs, its, x in outs[1:]:
        assert (segs, its) == outs[0][:2]
        np.testing.assert_array_equal(x, outs[0][2])


def test_rr_rejects_bad_arguments(ctx):
    rp, ci, v = P.poisson2d(12)
    A = ctx.csr(rp, ci, v)
    other = ctx.csr(rp, ci, v)
    A32 = A.convert(fsx.F32)
    A32.set_values_sym_scaled(A)
    b, x = ctx.vec(P.uniform(144, 1)), ctx.zeros(144)
    good = fsx.Settings(maxiter=100, rtol=1e-2)
    for bad in (fsx.Settings(maxiter=100, rtol=1.0), fsx.Settings(maxiter=100, rtol=0.0),
                fsx.Settings(maxiter=0, rtol=1e-2)):
        with pytest.raises(fsx.FsxError):
            fsx.ir_cg_rr_solve(A, A32, b, x, 1e-10, bad)
    with pytest.raises(fsx.FsxError):
        fsx.ir_cg_rr_solve(A, other.convert(fsx.F32), b, x, 1e-10, good)
    with pytest.raises(fsx.FsxError):
        fsx.ir_cg_rr_solve(A, A32, b, x, 1e-10, good, max_segments=-1)
    with pytest.raises(fsx.FsxError):
        fsx.ir_cg_rr_solve(A, A32, b, b, 1e-10, good)
    info = fsx.ir_cg_rr_solve(A, A32, b, x, 1e-10, good)
    assert info.status == 1 and info.rel_res_final <= 1e-10
