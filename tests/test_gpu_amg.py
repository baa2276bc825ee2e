"""Device UA-AMG (fsx_amg_*, mg::ua::bound_solver) against the oracle (gpu).

Bar: the hierarchy is identical to the oracle's -- level sizes, the
aggregate of every row and every coarse matrix entry bit for bit; V,
relaxed-W and relaxed-kappa cycles bit-identical to the oracle's
sequential cycles (smoothers, residuals and grid transfers repeat the
reference's arithmetic in its order; the coarsest solve is the same LU);
the K-cycle, whose dot products are device reductions, within 1e-11;
AMG-preconditioned CG reaches 1e-10 in far fewer iterations than Jacobi CG.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from flecsolve_amd import fsx
from flecsolve_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def lognormal3d(n, seed, itype=np.int32):
    rp, ci, v = P.diffusion3d(n, P.lognormal_field(n, 1.0, 3.0, seed))
    return rp.astype(itype), ci.astype(itype), v


def pair(ctx, rp, ci, v, **kw):
    s = dict(max_levels=8, maxiter=2, atol=0.0)
    s.update(kw)
    A = ctx.csr(rp, ci, v)
    M = fsx.AMG(A, fsx.AmgSettings(**s))
    Mo = O.Amg(O.Csr(rp, ci, v), O.AmgSettings(**s))
    return A, M, Mo


@pytest.mark.parametrize("case,passes", [("d3d", 2), ("d3d", 3), ("p2d64", 2)])
def test_hierarchy_bit_identical(ctx, case, passes):
    if case == "d3d":
        rp, ci, v = lognormal3d(12, 21)
    else:
        rp, ci, v = P.poisson2d(64)
        rp, ci = rp.astype(np.int64), ci.astype(np.int64)
    A, M, Mo = pair(ctx, rp, ci, v, pairwise_passes=passes)
    lv = M.levels()
    assert len(lv) == Mo.nlevels() >= 3
    for l, (rows, nnz) in enumerate(lv):
        orp, oci, ov = Mo.level(l)
        assert rows == orp.size - 1 and nnz == oci.size
        if l + 1 < len(lv):
            np.testing.assert_array_equal(M.aggregates(l), Mo.aggregates(l))
        if l > 0:
            grp, gci, gv = M.level_matrix(l).download()
            np.testing.assert_array_equal(grp, orp)
            np.testing.assert_array_equal(gci, oci)
            np.testing.assert_array_equal(gv, ov)


@pytest.mark.parametrize("cycle", [0, 1, 2])
@pytest.mark.parametrize("nrelax,smoother", [(1, 0), (2, 0), (2, 1)])
def test_cycles_bit_identical(ctx, cycle, nrelax, smoother):
    rp, ci, v = lognormal3d(14, 31)
    A, M, Mo = pair(ctx, rp, ci, v, cycle=cycle, nrelax=nrelax, kappa=2, maxiter=3,
                    smoother=smoother)
    b = P.uniform(rp.size - 1, 32)
    x = ctx.zeros(b.size)
    cycles, res = M.apply(ctx.vec(b), x)
    xo, co, reso = Mo.apply(b)
    assert cycles == co == 3
    np.testing.assert_array_equal(x.get(), xo)
    assert abs(res - reso) <= 1e-13 * reso


def test_kcycle_within_rounding(ctx):
    rp, ci, v = lognormal3d(14, 41)
    A, M, Mo = pair(ctx, rp, ci, v, cycle=3, kappa=2, maxiter=3)
    b = P.uniform(rp.size - 1, 42)
    x = ctx.zeros(b.size)
    M.apply(ctx.vec(b), x)
    xo, _, _ = Mo.apply(b)
    np.testing.assert_allclose(x.get(), xo, rtol=1e-11, atol=1e-13)


def test_atol_stops_cycles(ctx):
    rp, ci, v = P.poisson2d(48)
    A, M, Mo = pair(ctx, rp, ci, v, maxiter=60, atol=1e-4)
    b = P.uniform(48 * 48, 5)
    x = ctx.zeros(b.size)
    cycles, res = M.apply(ctx.vec(b), x)
    xo, co, reso = Mo.apply(b)
    assert cycles == co < 60 and res <= 1e-4
    np.testing.assert_array_equal(x.get(), xo)


def test_amg_preconditioned_cg(ctx):
    n = 32
    rp, ci, v = lognormal3d(n, 51)
    N = n ** 3
    A = ctx.csr(rp, ci, v)
    M = fsx.AMG(A, fsx.AmgSettings(max_levels=10, maxiter=1, atol=0.0))
    b = P.uniform(N, 52)
    bd = ctx.vec(b)
    S = sp.csr_matrix((v, ci, rp))
    s = fsx.Settings(maxiter=500, rtol=1e-10, use_zero_guess=True)
    x = ctx.zeros(N)
    info = M.cg(bd, x, s)
    rel = np.linalg.norm(b - S @ x.get()) / np.linalg.norm(b)
    xj = ctx.zeros(N)
    ij = fsx.cg_solve(A, bd, xj, s, dinv=A.diag_inv())
    print(f"AMG-CG {info.iters} iterations, Jacobi-CG {ij.iters}; true rel res {rel:.2e}")
    assert info.reason == "converged_rtol" and rel <= 1.1e-10
    assert info.iters * 3 < ij.iters
    xf = ctx.zeros(N)
    inf = M.cg(bd, xf, s, flexible=True)
    assert inf.reason == "converged_rtol"


def test_errors(ctx):
    rp, ci, v = P.poisson2d(8)
    A = ctx.csr(rp, ci, v)
    with pytest.raises(fsx.FsxError):
        fsx.AMG(A, fsx.AmgSettings(pairwise_passes=4))
    with pytest.raises(fsx.FsxError):
        fsx.AMG(A, fsx.AmgSettings(cycle=7))
    big = ctx.csr(*P.poisson2d(80))
    with pytest.raises(fsx.FsxError):  # one level: a 6400-row dense LU
        fsx.AMG(big, fsx.AmgSettings(max_levels=1))
    M = fsx.AMG(A)
    with pytest.raises(fsx.FsxError):
        M.apply(ctx.vec(np.ones(63)), ctx.zeros(64))
