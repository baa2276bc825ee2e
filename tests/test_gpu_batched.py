"""Batched persistent CG (fsx_batch_*) against the oracle, system by system (gpu).

Each system is an independent lognormal-k diffusion problem; the oracle solves
each one separately with the reference's serial CG + Dinv.  Bar: every system
converged, iteration counts within +-2, x within 1e-10 relative.
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


def make(nsys=24, emin=3, emax=12, sigma=1.0, seed=4001):
    e, rows, rp, ci, v = P.batch(nsys, emin, emax, seed, 4100, sigma)
    b = P.uniform(int(rows[-1]), 4200)
    return e, rows, rp, ci, v, b


def local(rows, rp, ci, v, s):
    lo, hi = rows[s], rows[s + 1]
    e0, e1 = rp[lo], rp[hi]
    return (rp[lo:hi + 1] - e0).astype(np.int64), (ci[e0:e1] - lo).astype(np.int64), v[e0:e1]


def check_against_oracle(rows, rp, ci, v, b, xg, it, st, rtol=1e-10, slack=2):
    for s in range(rows.size - 1):
        sl = slice(rows[s], rows[s + 1])
        Ao = O.Csr(*local(rows, rp, ci, v, s))
        xo, io, _ = O.cg(Ao, b[sl], np.zeros(rows[s + 1] - rows[s]),
                         O.Settings(maxiter=1000, rtol=rtol, use_zero_guess=True),
                         P=O.prec(1, Ao.dinv()))
        assert st[s] == 1, (s, st[s])
        assert abs(int(it[s]) - io.iters) <= slack, (s, it[s], io.iters)
        assert rel(xg[sl], xo) <= 1e-10, s


@pytest.mark.parametrize("cfg", [(256, 2, 1), (1024, 2, 1), (256, 0, 1), (1024, 0, 1),
                                 (64, 0, 1), (256, 1, 2), (512, 1, 4), (256, 1, 8)])
def test_batched_cg_matches_oracle(ctx, cfg):
    e, rows, rp, ci, v, b = make()
    A = ctx.csr(rp, ci, v)
    assert A.plan() == (3, 0)
    A.set_plan(2)   # also build the sliced layout for mapping 0
    A.set_plan(3)
    B = fsx.Batch(A, rows)
    assert B.config()[1] == 2
    B.set_config(*cfg)
    assert B.config() == cfg
    d = A.diag_inv()
    bd, x = ctx.vec(b), ctx.zeros(b.size)
    it, st, rs = B.cg(bd, x, fsx.Settings(maxiter=1000, rtol=1e-10, use_zero_guess=True),
                      dinv=d)
    check_against_oracle(rows, rp, ci, v, b, x.get(), it, st)


def test_batched_same_result_any_order_or_config(ctx):
    # a system's arithmetic is independent of which workgroup solved it
    e, rows, rp, ci, v, b = make(40, 4, 10)
    A = ctx.csr(rp, ci, v)
    B = fsx.Batch(A, rows)
    d = A.diag_inv()
    bd = ctx.vec(b)
    outs = []
    for wg in (256, 1024):
        B.set_config(wg, 2, 1)
        x = ctx.zeros(b.size)
        it, st, rs = B.cg(bd, x, fsx.Settings(maxiter=500, rtol=1e-10, use_zero_guess=True),
                          dinv=d)
        outs.append((x.get(), it))
    # workgroup size changes the reduction tree, so compare to rounding
    assert np.abs(outs[0][1] - outs[1][1]).max() <= 2
    assert rel(outs[0][0], outs[1][0]) <= 1e-10
    # repeat with the same config: bit-identical
    x = ctx.zeros(b.size)
    it, st, rs = B.cg(bd, x, fsx.Settings(maxiter=500, rtol=1e-10, use_zero_guess=True),
                      dinv=d)
    np.testing.assert_array_equal(x.get(), outs[1][0])


def test_batched_autotune(ctx):
    e, rows, rp, ci, v, b = make(64, 4, 16, seed=77)
    A = ctx.csr(rp, ci, v)
    B = fsx.Batch(A, rows)
    d = A.diag_inv()
    bd, x = ctx.vec(b), ctx.zeros(b.size)
    ms = B.tune(bd, x, d, sample=16, iters=10)
    assert ms > 0
    wg, mapping, lanes = B.config()
    assert wg in (256, 512, 1024) and mapping in (1, 2) and lanes in (1, 2, 4, 8)
    x = ctx.zeros(b.size)
    it, st, rs = B.cg(bd, x, fsx.Settings(maxiter=1000, rtol=1e-10, use_zero_guess=True),
                      dinv=d)
    check_against_oracle(rows, rp, ci, v, b, x.get(), it, st)


@pytest.mark.parametrize("jacobi", [1, 2])
@pytest.mark.parametrize("sigma", [1.0, 2.5])
def test_batched_ir_reaches_1e10(ctx, sigma, jacobi):
    e, rows, rp, ci, v, b = make(32, 4, 14, sigma=sigma)
    A = ctx.csr(rp, ci, v)
    A32 = A.convert(fsx.F32)
    if jacobi == 2:
        A32.set_values_sym_scaled(A)
    B = fsx.Batch(A, rows)
    bd, x = ctx.vec(b), ctx.zeros(b.size)
    outer, inner, st, rs = B.ir_cg(A32, bd, x, 1e-10, fsx.Settings(maxiter=2000, rtol=1e-4),
                                   jacobi=jacobi)
    assert (st == 1).all()
    S = sp.csr_matrix((v, ci, rp))
    xg = x.get()
    for s in range(rows.size - 1):
        sl = slice(rows[s], rows[s + 1])
        r = b[sl] - (S[sl] @ xg)
        assert np.linalg.norm(r) / np.linalg.norm(b[sl]) <= 1.01e-10, s
        assert outer[s] >= 1 and inner[s] >= outer[s]


@pytest.mark.parametrize("sigma", [1.0, 2.5])
def test_batched_rr_against_model(ctx, sigma):
    """Per system, the FP32 CG with FP64 residual replacement
    (fsx_batched_rr_cg): every system at true residual <= 1e-10, each
    system's segment and FP32 iteration counts those of the CPU model in its
    two-reduction form with FP64 row sums (tests/rr_model.py; +-1 / +-3; the
    model with FP32 row sums is 5-7 iterations away on some systems, so the
    bar separates the two products), fewer FP32 iterations
    in total than the restarted refinement on the same FP32 matrix, and the
    same x bit for bit under the sliced and slice-offset mappings (both sum
    rows serially)."""
    import rr_model as M
    e, rows, rp, ci, v, b = make(32, 4, 14, sigma=sigma)
    A = ctx.csr(rp, ci, v)
    A.set_plan(2)   # the sliced layout too, for mapping 0
    A.set_plan(3)
    A32 = A.convert(fsx.F32)
    A32.set_values_sym_scaled(A)
    B = fsx.Batch(A, rows)
    bd = ctx.vec(b)
    inner = fsx.Settings(maxiter=2000, rtol=1e-2)
    outs = []
    # the same workgroup size (it sets the reduction tree), two row mappings
    for cfg in ((256, 2, 1), (256, 0, 1)):
        B.set_config(*cfg)
        x = ctx.zeros(b.size)
        outs.append((B.rr_cg(A32, bd, x, 1e-10, inner), x.get()))
    (seg, its, st, rs), xg = outs[0]
    np.testing.assert_array_equal(outs[1][1], xg)
    np.testing.assert_array_equal(outs[1][0][1], its)
    assert (st == 1).all()
    S = sp.csr_matrix((v, ci, rp))
    for k in range(rows.size - 1):
        sl = slice(rows[k], rows[k + 1])
        assert np.linalg.norm(b[sl] - S[sl] @ xg) / np.linalg.norm(b[sl]) <= 1.01e-10, k
        assert abs(rs[k] - np.linalg.norm(b[sl] - S[sl] @ xg)) <= 1e-3 * rs[k] + 1e-300
        _, msegs, mits, mst, _ = M.rr_solve(*local(rows, rp, ci, v, k), b[sl], 1e-10, 1e-2,
                                            2000, 100, single_reduction=False, acc64=True)
        assert mst == 1
        assert abs(int(seg[k]) - msegs) <= 1, (k, seg[k], msegs)
        assert abs(int(its[k]) - mits) <= 3, (k, its[k], mits)
    x2 = ctx.zeros(b.size)
    outer, inner_r, st2, _ = B.ir_cg(A32, bd, x2, 1e-10, fsx.Settings(maxiter=2000, rtol=1e-4),
                                     jacobi=2)
    assert (st2 == 1).all()
    assert its.sum() < inner_r.sum()
    with pytest.raises(fsx.FsxError):
        B.rr_cg(A32, bd, x, 1e-10, fsx.Settings(maxiter=2000, rtol=1.0))


def test_batched_edge_cases(ctx):
    e, rows, rp, ci, v, b = make(6, 2, 5)
    b[rows[2]:rows[3]] = 0.0       # zero right-hand side: converged at once
    A = ctx.csr(rp, ci, v)
    B = fsx.Batch(A, rows)
    bd, x = ctx.vec(b), ctx.zeros(b.size)
    it, st, rs = B.cg(bd, x, fsx.Settings(maxiter=200, rtol=1e-10, use_zero_guess=True),
                      dinv=A.diag_inv())
    assert it[2] == 0 and st[2] == 1
    assert (st == 1).all()
    # maxiter exhausted: iters 0, diverged_iters (cg.hh:135-136)
    it, st, rs = B.cg(bd, x, fsx.Settings(maxiter=1, rtol=1e-14, use_zero_guess=True))
    assert ((it == 0) & (st == 4) | (it == 0) & (st == 1)).all()
    # coupled systems are rejected
    rp2, ci2 = rp.copy(), ci.copy()
    ci2[rp2[rows[1]]] = rows[0]     # first row of system 1 points into system 0
    A2 = ctx.csr(rp2, ci2, v)
    with pytest.raises(fsx.FsxError):
        fsx.Batch(A2, rows)


def test_batched_full_edge_range_matches_oracle(ctx):
    """Edges spanning 8..48 as the 2048-system workload draws them, with an
    8^3 (512 rows) and a 48^3 system (110,592 rows in one workgroup): FP64
    CG + Jacobi under the autotuned mapping and under each mapping family,
    system by system against the oracle; then the FP32-inner refinement,
    every system at true residual <= 1e-10 and within 1e-8 of the oracle's
    1e-10 CG solution (each 1e-10-residual solution of these systems carries
    a forward error of up to 5e-10: the oracle's against its own rtol 1e-14
    solve)."""
    e, rows, rp, ci, v = P.batch(12, 8, 48, 4007, 4100, 1.0)
    assert 8 in e and 48 in e
    b = P.uniform(int(rows[-1]), 4200)
    A = ctx.csr(rp, ci, v)
    assert A.plan() == (3, 0)
    A.set_plan(2)   # the sliced layout too, for mapping 0
    A.set_plan(3)
    B = fsx.Batch(A, rows)
    d = A.diag_inv()
    bd = ctx.vec(b)
    s64 = fsx.Settings(maxiter=1000, rtol=1e-10, use_zero_guess=True)
    x = ctx.zeros(b.size)
    B.tune(bd, x, d, sample=12, iters=10)
    for cfg in (None, (1024, 2, 1), (256, 0, 1), (512, 1, 4)):
        if cfg is not None:
            B.set_config(*cfg)
        x = ctx.zeros(b.size)
        it, st, rs = B.cg(bd, x, s64, dinv=d)
        check_against_oracle(rows, rp, ci, v, b, x.get(), it, st)
    A32 = A.convert(fsx.F32)
    A32.set_values_sym_scaled(A)
    x = ctx.zeros(b.size)
    outer, inner, st, rs = B.ir_cg(A32, bd, x, 1e-10, fsx.Settings(maxiter=2000, rtol=1e-4),
                                   jacobi=2)
    assert (st == 1).all()
    S = sp.csr_matrix((v, ci, rp))
    xg = x.get()
    for k in range(rows.size - 1):
        sl = slice(rows[k], rows[k + 1])
        assert np.linalg.norm(b[sl] - S[sl] @ xg) / np.linalg.norm(b[sl]) <= 1.01e-10, k
        Ao = O.Csr(*local(rows, rp, ci, v, k))
        xo, io, _ = O.cg(Ao, b[sl], np.zeros(rows[k + 1] - rows[k]),
                         O.Settings(maxiter=1000, rtol=1e-10, use_zero_guess=True),
                         P=O.prec(1, Ao.dinv()))
        assert rel(xg[sl], xo) <= 1e-8, (k, e[k])
