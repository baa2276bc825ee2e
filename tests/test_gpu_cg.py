"""Fused, graph-captured CG (fsx_cg_solve) and FP32-inner / FP64-outer
refinement against the oracle (gpu).

Bar (from the hot-path contract): iteration counts within +-2 of the
oracle's, solutions within 1e-10 relative of the oracle's.  The reductions
sum in a different (tree) order than the reference's serial loop, so the two
trajectories drift apart at the level of rounding; the solution comparison
uses a tolerance a factor below the solve tolerance where the problem's
conditioning allows, and otherwise compares against the exact solution.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from flecsolve_amd import fsx
from flecsolve_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def poisson(n):
    rp, ci, v = P.poisson2d(n)
    return rp, ci, v, O.Csr(rp, ci, v)


def diffusion(n, sigma=1.0, seed=2001):
    k = P.lognormal_field(n, sigma, 8.0, seed)
    rp, ci, v = P.diffusion3d(n, k)
    return rp, ci, v, O.Csr(rp, ci, v)


def test_cg_poisson64_cfg1(ctx):
    """2D Poisson 64x64, unpreconditioned CG to rtol 1e-8, x0 = 0."""
    rp, ci, v, Ao = poisson(64)
    b = P.uniform(64 * 64, P.seeds()["poisson2d_rhs"])
    s = dict(maxiter=1000, rtol=1e-8, use_zero_guess=True)
    xo, io, ho = O.cg(Ao, b, np.zeros_like(b), O.Settings(**s), hist_cap=1000)
    A = ctx.csr(rp, ci, v)
    bd, xd = ctx.vec(b), ctx.zeros(b.size)
    info = fsx.cg_solve(A, bd, xd, fsx.Settings(**s))
    assert info.reason == io.reason == "converged_rtol"
    assert abs(info.iters - io.iters) <= 2
    assert rel(xd.get(), xo) <= 1e-10
    h = fsx.cg_history(ctx)
    m = min(len(h), io.iters)
    np.testing.assert_allclose(h[:m], ho[:m], rtol=1e-6)
    assert info.rhs_norm == io.rhs_norm


@pytest.mark.parametrize("n", [16, 48])
def test_cg_jacobi_diffusion(ctx, n):
    rp, ci, v, Ao = diffusion(n)
    b = P.uniform(n ** 3, 2002)
    s = dict(maxiter=3000, rtol=1e-10, use_zero_guess=True)
    xo, io, _ = O.cg(Ao, b, np.zeros_like(b), O.Settings(**s), P=O.prec(1, Ao.dinv()))
    A = ctx.csr(rp, ci, v)
    d = A.diag_inv()
    xd = ctx.zeros(b.size)
    info = fsx.cg_solve(A, ctx.vec(b), xd, fsx.Settings(**s), dinv=d)
    assert info.reason == "converged_rtol"
    assert abs(info.iters - io.iters) <= 2
    assert rel(xd.get(), xo) <= 1e-10
    S = sp.csr_matrix((v, ci, rp))
    assert np.linalg.norm(b - S @ xd.get()) / np.linalg.norm(b) < 2e-10


def test_cg_nonzero_guess_and_repeat(ctx):
    rp, ci, v, Ao = diffusion(20)
    n = 20 ** 3
    b = P.uniform(n, 1)
    x0 = O.set_random(n, 7)
    s = dict(maxiter=500, rtol=1e-9)
    xo, io, _ = O.cg(Ao, b, x0, O.Settings(**s))
    A = ctx.csr(rp, ci, v)
    for _ in range(2):  # second call replays the cached graph
        xd = ctx.vec(x0)
        info = fsx.cg_solve(A, ctx.vec(b), xd, fsx.Settings(**s))
        assert abs(info.iters - io.iters) <= 2
        assert rel(xd.get(), xo) <= 1e-9
        assert info.sol_norm_initial == io.sol_norm_initial


def test_cg_quirks(ctx):
    rp, ci, v, Ao = poisson(8)
    A = ctx.csr(rp, ci, v)
    b = P.uniform(64, 1)
    # maxiter hit: iters == 0, diverged_iters (cg.hh:135-136)
    info = fsx.cg_solve(A, ctx.vec(b), ctx.zeros(64),
                        fsx.Settings(maxiter=3, rtol=1e-12, use_zero_guess=True))
    assert info.iters == 0 and info.reason == "diverged_iters"
    xo, io, _ = O.cg(Ao, b, np.zeros(64), O.Settings(maxiter=3, rtol=1e-12,
                                                    use_zero_guess=True))
    assert info.res_norm_final == pytest.approx(io.res_norm_final, rel=1e-6)
    # zero rhs: converged before the loop, b_norm := 1
    info = fsx.cg_solve(A, ctx.zeros(64), ctx.zeros(64),
                        fsx.Settings(maxiter=3, rtol=1e-12, use_zero_guess=True))
    assert info.iters == 0 and info.reason == "converged_rtol" and info.rhs_norm == 1.0
    # maxiter == 0
    info = fsx.cg_solve(A, ctx.vec(b), ctx.zeros(64),
                        fsx.Settings(maxiter=0, rtol=1e-12, use_zero_guess=True))
    assert info.reason == "diverged_iters"


def test_cg_chunk_boundaries(ctx):
    # convergence exactly at, before and after a captured chunk boundary
    d = np.repeat(np.arange(1.0, 18.0), 3)  # 17 distinct eigenvalues
    n = d.size
    rp = np.arange(n + 1, dtype=np.int32)
    A = ctx.csr(rp, np.arange(n, dtype=np.int32), d)
    Ao = O.Csr(rp, np.arange(n, dtype=np.int32), d)
    b = O.set_random(n, 3)
    for maxiter in (15, 16, 17, 40):
        s = dict(maxiter=maxiter, rtol=1e-12, use_zero_guess=True)
        xo, io, _ = O.cg(Ao, b, np.zeros(n), O.Settings(**s))
        xd = ctx.zeros(n)
        info = fsx.cg_solve(A, ctx.vec(b), xd, fsx.Settings(**s))
        assert (info.iters, info.reason) == (io.iters, io.reason), maxiter
        np.testing.assert_allclose(xd.get(), xo, rtol=1e-10, atol=1e-14)


def test_fp32_cg(ctx):
    rp, ci, v, Ao = diffusion(24)
    n = 24 ** 3
    b = P.uniform(n, 5).astype(np.float32)
    A32 = ctx.csr(rp, ci, v).convert(fsx.F32)
    d = ctx.csr(rp, ci, v).diag_inv(fsx.F32)
    xd = ctx.zeros(n, fsx.F32)
    info = fsx.cg_solve(A32, ctx.vec(b), xd, fsx.Settings(maxiter=500, rtol=1e-4,
                                                         use_zero_guess=True), dinv=d)
    assert info.reason == "converged_rtol"
    S = sp.csr_matrix((v, ci, rp))
    x = xd.get().astype(np.float64)
    assert np.linalg.norm(b - S @ x) / np.linalg.norm(b) < 2e-4


@pytest.mark.parametrize("jacobi", [1, 2])
@pytest.mark.parametrize("sigma", [1.0, 2.5])
def test_ir_reaches_1e10_and_matches_fp64(ctx, sigma, jacobi):
    n = 32
    rp, ci, v, Ao = diffusion(n, sigma, seed=3001)
    N = n ** 3
    b = P.uniform(N, 3002)
    A64 = ctx.csr(rp, ci, v)
    A32 = A64.convert(fsx.F32)
    if jacobi == 2:
        A32.set_values_sym_scaled(A64)
    xd = ctx.zeros(N)
    info = fsx.ir_cg_solve(A64, A32, ctx.vec(b), xd, 1e-10,
                           fsx.Settings(maxiter=2000, rtol=1e-4), jacobi=jacobi)
    assert info.status == 1 and info.rel_res_final <= 1e-10
    S = sp.csr_matrix((v, ci, rp))
    x = xd.get()
    assert np.linalg.norm(b - S @ x) / np.linalg.norm(b) <= 1.01e-10
    xo, io, _ = O.cg(Ao, b, np.zeros(N), O.Settings(maxiter=5000, rtol=1e-12,
                                                   use_zero_guess=True),
                     P=O.prec(1, Ao.dinv()))
    # both are 1e-10-residual solutions: error bounded by cond * 1e-10
    assert rel(x, xo) < 1e-6


@pytest.mark.parametrize("plan", [4, 3, 2])
def test_grid_stride_sizes_iteration_matched(ctx, plan):
    """More rows than one capped SpMV grid covers (> 8192 x 256), so each
    thread owns several rows: SpMV+dot and 25 CG iterations against the
    oracle, iteration-matched (both stop at maxiter), on the symmetric
    stencil, slice-offset and sliced layouts."""
    n = 132
    rp, ci, v, Ao = diffusion(n)
    N = n ** 3
    assert N > 8192 * 256
    b = P.uniform(N, 2002)
    A = ctx.csr(rp, ci, v)
    assert A.plan() == (4, 0)
    A.set_plan(plan)
    bd = ctx.vec(b)
    w = ctx.zeros(N)
    pw = fsx.spmv_dot(A, bd, w)
    wref = Ao.spmv(b)
    np.testing.assert_array_equal(w.get(), wref)
    assert abs(pw - O.dot(wref, b)) <= 1e-12 * np.abs(wref * b).sum()
    s = dict(maxiter=25, rtol=1e-12, use_zero_guess=True)
    xo, io, ho = O.cg(Ao, b, np.zeros(N), O.Settings(**s), P=O.prec(1, Ao.dinv()),
                      hist_cap=25)
    xd = ctx.zeros(N)
    info = fsx.cg_solve(A, bd, xd, fsx.Settings(**s), dinv=A.diag_inv())
    assert info.reason == io.reason == "diverged_iters"
    np.testing.assert_allclose(fsx.cg_history(ctx), ho, rtol=1e-10)
    assert rel(xd.get(), xo) <= 1e-12


def test_wide_block_shape_iteration_matched(ctx):
    """From 2^23 rows the FP64 reducing product runs in 512-thread blocks
    (another block partial tree than below that size): 204^3 = 8,489,664
    rows, the SpMV+dot bit-compared and 25 CG + Jacobi iterations against
    the oracle, iteration-matched (both stop at maxiter)."""
    n = 204
    rp, ci, v, Ao = diffusion(n)
    N = n ** 3
    assert N >= 1 << 23
    b = P.uniform(N, 2002)
    A = ctx.csr(rp, ci, v)
    assert A.plan() == (4, 0)
    bd = ctx.vec(b)
    w = ctx.zeros(N)
    pw = fsx.spmv_dot(A, bd, w)
    wref = Ao.spmv(b)
    np.testing.assert_array_equal(w.get(), wref)
    assert abs(pw - O.dot(wref, b)) <= 1e-12 * np.abs(wref * b).sum()
    s = dict(maxiter=25, rtol=1e-12, use_zero_guess=True)
    xo, io, ho = O.cg(Ao, b, np.zeros(N), O.Settings(**s), P=O.prec(1, Ao.dinv()),
                      hist_cap=25)
    xd = ctx.zeros(N)
    info = fsx.cg_solve(A, bd, xd, fsx.Settings(**s), dinv=A.diag_inv())
    assert info.reason == io.reason == "diverged_iters"
    np.testing.assert_allclose(fsx.cg_history(ctx), ho, rtol=1e-10)
    assert rel(xd.get(), xo) <= 1e-12


def test_layouts_give_identical_trajectories(ctx):
    """Every serial-row layout gives the same SpMV bits; the fused dot
    products are reduced over each layout's own block grid, so iterates
    agree to rounding and the iteration counts match."""
    n = 40
    rp, ci, v, Ao = diffusion(n)
    b = P.uniform(n ** 3, 77)
    A = ctx.csr(rp, ci, v)
    d = A.diag_inv()
    bd = ctx.vec(b)
    out = []
    for plan in (4, 3, 2, 0):
        A.set_plan(plan)
        x = ctx.zeros(n ** 3)
        info = fsx.cg_solve(A, bd, x, fsx.Settings(maxiter=2000, rtol=1e-10,
                                                   use_zero_guess=True), dinv=d)
        out.append((info.iters, x.get()))
    for it, x in out[1:]:
        assert abs(it - out[0][0]) <= 1
        assert rel(x, out[0][1]) <= 1e-12


def test_graphs_after_matrix_rebuild(ctx):
    """Captured solver graphs hold the device arrays of the matrix they were
    captured for.  Matrices of one shape are built, solved and freed in turn
    on one context (their buffers may come back at the old addresses), and
    every solve must match its own residual -- no graph replays against the
    arrays of a freed matrix."""
    import gc
    n = 24
    N = n ** 3
    for rep, sigma in enumerate((1.0, 2.0, 1.5)):
        rp, ci, v, Ao = diffusion(n, sigma, seed=4000 + rep)
        b = P.uniform(N, 4100 + rep)
        A = ctx.csr(rp, ci, v)
        A32 = A.convert(fsx.F32)
        S = sp.csr_matrix((v, ci, rp))
        x = ctx.zeros(N)
        info = fsx.ir_cg_solve(A, A32, ctx.vec(b), x, 1e-10,
                               fsx.Settings(maxiter=2000, rtol=1e-4))
        assert info.status == 1
        assert np.linalg.norm(b - S @ x.get()) / np.linalg.norm(b) <= 1.01e-10
        x2 = ctx.zeros(N)
        i2 = fsx.cg_solve(A, ctx.vec(b), x2, fsx.Settings(maxiter=3000, rtol=1e-9,
                                                          use_zero_guess=True),
                          dinv=A.diag_inv())
        assert i2.reason == "converged_rtol"
        assert np.linalg.norm(b - S @ x2.get()) / np.linalg.norm(b) <= 1e-9
        del A, A32, x, x2
        gc.collect()


@pytest.mark.parametrize("plan", [4, 3, 2, 0])
def test_sym_scaled_f32_values(ctx, plan):
    """S A S in FP32, S = 1/sqrt(diag): values bit for bit against numpy
    (a_ij (s_i s_j) rounded once, so symmetric A gives bitwise symmetric
    values and keeps the symmetric stencil layout), on each SpMV layout;
    errors for a missing or non-positive diagonal."""
    n = 20
    rp, ci, v, _ = diffusion(n, 2.0, seed=5001)
    A64 = ctx.csr(rp, ci, v)
    A64.set_plan(plan)
    A32 = A64.convert(fsx.F32)
    A32.set_plan(plan)
    A32.set_values_sym_scaled(A64)
    assert A32.plan() == (plan, 0)
    rows = np.repeat(np.arange(n ** 3), np.diff(rp))
    d = v[ci == rows]
    sv = 1.0 / np.sqrt(d)
    want = (v * (sv[rows] * sv[ci])).astype(np.float32)
    np.testing.assert_array_equal(A32.download()[2], want)
    W = sp.csr_matrix((want, ci, rp))
    assert (W != W.T).nnz == 0
    x = P.uniform(n ** 3, 5002).astype(np.float32)
    y = ctx.zeros(n ** 3, fsx.F32)
    A32.spmv(ctx.vec(x, fsx.F32), y)
    ref = sp.csr_matrix((want.astype(np.float64), ci, rp)) @ x.astype(np.float64)
    assert np.abs(y.get() - ref).max() <= 1e-5 * np.abs(ref).max()
    bad = ctx.csr(np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32),
                  np.array([1.0, 1.0]))
    with pytest.raises(fsx.FsxError):
        bad.convert(fsx.F32).set_values_sym_scaled(bad)
    neg = ctx.csr(np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32),
                  np.array([1.0, -1.0]))
    with pytest.raises(fsx.FsxError):
        neg.convert(fsx.F32).set_values_sym_scaled(neg)


def test_new_entry_points_reject_bad_arguments(ctx):
    """Errors, not faults: a refinement (jacobi = 2) or refined adjoint whose
    FP32 matrix is not a copy of A's pattern, an unknown mode, a stage matrix
    that is F itself or does not share F's pattern."""
    rp, ci, v = P.poisson2d(12)
    A = ctx.csr(rp, ci, v)
    other = ctx.csr(rp, ci, v)          # same values, its own index arrays
    A32_other = other.convert(fsx.F32)
    b, x = ctx.vec(P.uniform(144, 1)), ctx.zeros(144)
    inner = fsx.Settings(maxiter=100, rtol=1e-4)
    with pytest.raises(fsx.FsxError):
        fsx.ir_cg_solve(A, A32_other, b, x, 1e-10, inner, jacobi=2)
    with pytest.raises(fsx.FsxError):
        fsx.ir_cg_solve(A, A.convert(fsx.F32), b, x, 1e-10, inner, jacobi=3)
    with pytest.raises(fsx.FsxError):
        fsx.be_forward_ir(A, A32_other, b, 2, fsx.Settings(maxiter=100, rtol=1e-10), inner)
    with pytest.raises(fsx.FsxError):
        A.shift_scale(A, 1.0, -0.5)
    with pytest.raises(fsx.FsxError):
        other.convert(fsx.F64).shift_scale(A, 1.0, -0.5)
    # the good path still works afterwards
    A32 = A.convert(fsx.F32)
    A32.set_values_sym_scaled(A)
    info = fsx.ir_cg_solve(A, A32, b, x, 1e-10, inner, jacobi=2)
    assert info.status == 1


def test_cg_user_diagnostic(ctx):
    """The fused CG with the reference's user diagnostic (cg.hh:112-116):
    called after every iteration with the updated x and |r|, before the
    tolerance test.  A diagnostic that never stops gives fsx_cg_solve's
    iterates bit for bit, its |r| values are the residual history; one that
    stops at iteration 9 ends the solve there with converged_user and the
    oracle's x (its diagnostic stopping at the same call); the reference
    test's A-norm diagnostic (solvers/test/cg.cc:16-54: monotone decrease
    and the Trefethen-Bau bound for b = 0, x0 = set_random) holds; an
    exception in the diagnostic aborts the solve."""
    n = 40
    rp, ci, v = P.poisson2d(n)
    Ao = O.Csr(rp, ci, v)
    A = ctx.csr(rp, ci, v)
    d = A.diag_inv()
    b = P.uniform(n * n, 31)
    s = dict(maxiter=2000, rtol=1e-10, use_zero_guess=True)
    x1 = ctx.zeros(n * n)
    i1 = fsx.cg_solve(A, ctx.vec(b), x1, fsx.Settings(**s), dinv=d)
    h1 = fsx.cg_history(ctx)
    seen = []
    x2 = ctx.zeros(n * n)
    i2 = fsx.cg_solve_diag(A, ctx.vec(b), x2, fsx.Settings(**s),
                           lambda x, res: seen.append(res) and False, dinv=d)
    assert i2.as_dict() == i1.as_dict()
    assert np.array_equal(x2.get(), x1.get())
    np.testing.assert_array_equal(np.array(seen), h1)

    # stop at the 9th call: converged_user, iters = 9
    calls = []
    x3 = ctx.zeros(n * n)
    i3 = fsx.cg_solve_diag(A, ctx.vec(b), x3, fsx.Settings(**s),
                           lambda x, res: calls.append(res) or len(calls) == 9, dinv=d)
    oc = []
    xo, io, _ = O.cg(Ao, b, np.zeros(n * n), O.Settings(**s), P=O.prec(1, Ao.dinv()),
                     diag=lambda x, res: oc.append(res) or len(oc) == 9)
    assert i3.reason == io.reason == "converged_user"
    assert i3.iters == io.iters == 9
    assert rel(x3.get(), xo) <= 1e-13
    np.testing.assert_allclose(calls, oc, rtol=1e-12)

    # the reference test: b = 0, x0 = random, the A-norm of the error (= x)
    # falls monotonically and below 2 ((sqrt k - 1)/(sqrt k + 1))^i |e0|_A
    lam = np.array([4 - 2 * np.cos(np.pi * i / (n + 1)) - 2 * np.cos(np.pi * j / (n + 1))
                    for i in range(1, n + 1) for j in range(1, n + 1)])
    kappa = lam.max() / lam.min()
    S = sp.csr_matrix((v, ci, rp))
    x0 = np.random.default_rng(7).random(n * n)
    anorm = lambda y: np.sqrt(y @ (S @ y))
    e0 = anorm(x0)
    errs = []
    xd = ctx.vec(x0)
    i4 = fsx.cg_solve_diag(A, ctx.zeros(n * n), xd,
                           fsx.Settings(maxiter=2000, rtol=1e-9, use_zero_guess=False),
                           lambda x, res: errs.append(anorm(x.get())) and False)
    assert i4.reason == "converged_rtol" and i4.iters == len(errs)
    frac = (np.sqrt(kappa) - 1) / (np.sqrt(kappa) + 1)
    bound = 2 * frac ** np.arange(1, len(errs) + 1) * e0
    assert np.all(np.diff(np.concatenate([[e0], errs])) < 0)
    assert np.all(np.array(errs) < bound)

    class Stop(RuntimeError):
        pass

    def bad(x, res):
        raise Stop("from the diagnostic")

    with pytest.raises(Stop):
        fsx.cg_solve_diag(A, ctx.vec(b), ctx.zeros(n * n), fsx.Settings(**s), bad, dinv=d)
    # the context still solves afterwards
    x5 = ctx.zeros(n * n)
    i5 = fsx.cg_solve(A, ctx.vec(b), x5, fsx.Settings(**s), dinv=d)
    assert i5.as_dict() == i1.as_dict()
