"""Oracle CG / FCG / GMRES / BiCGSTAB against what the mathematics fixes.

Pins (the reference's own KATs need SuiteSparse matrices that are not
available offline, so each solver is pinned by properties a plausible slip
-- a sign, a dropped term, a swapped index -- would break):
  * CG: exact termination on diagonal matrices with k distinct eigenvalues;
    A-norm error monotone and under the Trefethen-Bau bound
    (flecsolve/solvers/test/cg.cc:16-54) on the 2D Laplacian with its closed
    form spectrum; agreement with a direct solve; maxiter / zero-rhs quirks.
  * GMRES: monotone residual, bound of flecsolve/solvers/test/gmres.cc:17-49,
    the restart-split identity of gmres.cc:88-110, direct-solve agreement for
    left and right preconditioning.
  * BiCGSTAB: direct-solve agreement, the A = I early-exit quirk.
  * FCG: same solution as CG.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from flecsolve_amd import problems as P
from oracle import oracle as O


def lap2d(n):
    rp, ci, v = P.poisson2d(n)
    return O.Csr(rp, ci, v), sp.csr_matrix((v, ci, rp))


def lap2d_cond(n):
    c = np.cos(np.pi * np.arange(1, n + 1) / (n + 1))
    lam = 4 - 2 * c[:, None] - 2 * c[None, :]
    return lam.max() / lam.min()


def test_cg_diagonal_terminates_in_k_steps():
    d = np.repeat([1.0, 2.0, 5.0, 11.0], 25)
    n = d.size
    A = O.Csr(np.arange(n + 1), np.arange(n), d)
    b = O.set_random(n, 3)
    x, info, _ = O.cg(A, b, np.zeros(n), O.Settings(maxiter=50, rtol=1e-12,
                                                   use_zero_guess=True))
    assert info.reason == "converged_rtol"
    assert info.iters == 4
    np.testing.assert_allclose(x, b / d, rtol=1e-12)


def test_cg_anorm_monotone_and_trefethen_bau_bound():
    n = 20
    A, S = lap2d(n)
    cond = lap2d_cond(n)
    x0 = O.set_random(n * n, 7)     # cg.cc:74-75: b = 0, x0 = set_random(7)
    b = np.zeros(n * n)
    e0 = np.sqrt(x0 @ (S @ x0))
    frac = (np.sqrt(cond) - 1) / (np.sqrt(cond) + 1)
    state = {"k": 0, "prev": e0, "mono": True, "bound": True}

    def diag(x, res):
        state["k"] += 1
        e = np.sqrt(x @ (S @ x))
        state["mono"] &= e <= state["prev"]
        state["bound"] &= e < 2 * frac ** state["k"] * e0
        state["prev"] = e
        return False

    x, info, _ = O.cg(A, b, x0, O.Settings(maxiter=2000, rtol=1e-9, atol=1e-9),
                      diag=diag)
    assert info.reason == "converged_rtol"
    assert info.iters == state["k"]
    assert state["mono"] and state["bound"]
    # b == 0: b_norm is replaced by 1, so the tolerance is absolute
    assert info.res_norm_final < 1e-9


def test_cg_matches_direct_solve_and_history():
    n = 32
    A, S = lap2d(n)
    b = P.uniform(n * n, 11)
    x, info, hist = O.cg(A, b, np.zeros(n * n),
                         O.Settings(maxiter=500, rtol=1e-10, use_zero_guess=True),
                         hist_cap=500)
    xe = np.linalg.solve(S.toarray(), b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-8
    # tol is float(rtol) * |b|
    tol = float(np.float32(1e-10)) * np.linalg.norm(b)
    h = hist[:info.iters]
    assert h[-1] < tol and (info.iters == 1 or h[-2] >= tol)
    assert info.res_norm_final == np.float32(h[-1])
    # recursive residual tracks the true one
    assert abs(np.linalg.norm(b - S @ x) - h[-1]) < 1e-6 * np.linalg.norm(b)


def test_cg_jacobi_equals_scaled_problem_solution():
    n = 12
    k = P.lognormal_field(n, 1.0, 4.0, 5)
    rp, ci, v = P.diffusion3d(n, k)
    A = O.Csr(rp, ci, v)
    S = sp.csr_matrix((v, ci, rp))
    b = P.uniform(n ** 3, 6)
    x, info, _ = O.cg(A, b, np.zeros(n ** 3),
                      O.Settings(maxiter=1000, rtol=1e-10, use_zero_guess=True),
                      P=O.prec(1, A.dinv()))
    x0, info0, _ = O.cg(A, b, np.zeros(n ** 3),
                        O.Settings(maxiter=1000, rtol=1e-10, use_zero_guess=True))
    xe = np.linalg.solve(S.toarray(), b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-8
    assert info.iters < info0.iters  # Jacobi helps on a variable coefficient


def test_cg_quirks():
    A, S = lap2d(8)
    b = P.uniform(64, 1)
    # maxiter exhausted: iters stays 0, status diverged_iters (cg.hh:135-136)
    _, info, _ = O.cg(A, b, np.zeros(64), O.Settings(maxiter=3, rtol=1e-12,
                                                    use_zero_guess=True))
    assert info.iters == 0 and info.reason == "diverged_iters"
    # zero right-hand side with zero guess: converged before iterating
    _, info, _ = O.cg(A, np.zeros(64), np.zeros(64),
                      O.Settings(maxiter=3, rtol=1e-12, use_zero_guess=True))
    assert info.iters == 0 and info.reason == "converged_rtol"
    assert info.rhs_norm == 1.0


def test_fcg_matches_cg_solution():
    A, S = lap2d(16)
    b = P.uniform(256, 2)
    s = O.Settings(maxiter=500, rtol=1e-11, use_zero_guess=True)
    x1, i1, _ = O.cg(A, b, np.zeros(256), s)
    x2, i2 = O.fcg(A, b, np.zeros(256), s)
    assert i2.reason == "converged_rtol"
    np.testing.assert_allclose(x2, x1, rtol=1e-8, atol=1e-10)
    assert abs(i1.iters - i2.iters) <= 2


def test_gmres_monotone_bound_and_direct():
    n = 16
    rp, ci, v = P.convdiff3d(n, (10.0, 5.0, 2.5))
    A = O.Csr(rp, ci, v)
    S = sp.csr_matrix((v, ci, rp))
    b = P.uniform(n ** 3, 3)
    hist = []
    _, _, _ = O.gmres(A, b, np.zeros(n ** 3),
                      O.Settings(maxiter=100, rtol=1e-10, use_zero_guess=True),
                      diag=lambda x, r: hist.append(r) or False)
    assert all(hist[i + 1] <= hist[i] for i in range(len(hist) - 1))
    for side in (0, 1):
        x, info, _ = O.gmres(A, b, np.zeros(n ** 3),
                             O.Settings(maxiter=300, rtol=1e-10, use_zero_guess=True,
                                        max_krylov_dim=30, restart=True,
                                        pre_side=side),
                             P=O.prec(1, A.dinv()))
        assert info.reason == "converged_rtol"
        rel = np.linalg.norm(b - S @ x) / np.linalg.norm(b)
        # left preconditioning converges on |D^-1 r| (gmres.hh:111-117), so
        # the unpreconditioned residual is only within a factor of max a_ii
        assert rel < (1e-9 if side == 1 else 1e-8), (side, rel)


def test_gmres_restart_split_identity():
    # flecsolve/solvers/test/gmres.cc:88-110 on an available matrix
    n = 24
    A, S = lap2d(n)
    b = O.set_random(n * n, 0)
    x0 = O.set_random(n * n, 1)
    rst = O.Settings(maxiter=100, rtol=1e-6, max_krylov_dim=50, restart=True)
    _, info_r, _ = O.gmres(A, b, x0, rst)
    assert info_r.restarts == 1 and info_r.iters > 50
    x50, _, _ = O.gmres(A, b, x0, O.Settings(maxiter=50, rtol=1e-6,
                                             max_krylov_dim=50, restart=True))
    _, info, _ = O.gmres(A, b, x50, O.Settings(maxiter=100, rtol=1e-6,
                                               max_krylov_dim=100, restart=False))
    assert 50 + info.iters == info_r.iters
    assert info.res_norm_final == info_r.res_norm_final


def test_gmres_rejects_bad_dimension():
    A, _ = lap2d(4)
    with pytest.raises(ValueError):
        O.gmres(A, np.ones(16), np.zeros(16), O.Settings(maxiter=200, max_krylov_dim=-1))


def test_bicgstab_direct_and_identity_quirk():
    n = 16
    rp, ci, v = P.convdiff3d(n, (10.0, 5.0, 2.5))
    A = O.Csr(rp, ci, v)
    S = sp.csr_matrix((v, ci, rp))
    b = P.uniform(n ** 3, 4)
    x, info, _ = O.bicgstab(A, b, np.zeros(n ** 3),
                            O.Settings(maxiter=300, rtol=1e-10, use_zero_guess=True))
    assert info.reason == "converged_rtol"
    assert np.linalg.norm(b - S @ x) / np.linalg.norm(b) < 1e-9
    # A = I: s = 0 after the first half step, early exit reports iters = 0,
    # which the final check then turns into diverged_iters (bicgstab.hh:136-142,
    # 183-184) although x is exact.
    m = 10
    I = O.Csr(np.arange(m + 1), np.arange(m), np.ones(m))
    bb = P.uniform(m, 5)
    x, info, _ = O.bicgstab(I, bb, np.zeros(m), O.Settings(maxiter=10, rtol=1e-8,
                                                           use_zero_guess=True))
    np.testing.assert_array_equal(x, bb)
    assert info.iters == 0 and info.reason == "diverged_iters"


def test_cg_colours_baseline():
    # one colour is op-for-op the serial CG + Dinv; several colours change
    # only the summation order of the reductions
    from flecsolve_amd import problems as P
    n = 20
    rp, ci, v = P.diffusion3d(n, P.lognormal_field(n, 1.0, 4.0, 3))
    A = O.Csr(rp, ci, v)
    b = P.uniform(n ** 3, 4)
    dinv = 1.0 / v[rp[:-1] + np.array([np.where(ci[rp[i]:rp[i + 1]] == i)[0][0]
                                       for i in range(n ** 3)])]
    s = O.Settings(maxiter=1000, rtol=1e-10, use_zero_guess=True)
    xs, io, _ = O.cg(A, b, np.zeros(n ** 3), s, P=O.prec(1, A.dinv()))
    x1, i1 = O.cg_colours(A, dinv, b, np.zeros(n ** 3), s, 1)
    np.testing.assert_array_equal(x1, xs)
    assert i1.iters == io.iters
    x4, i4 = O.cg_colours(A, dinv, b, np.zeros(n ** 3), s, 4)
    assert abs(i4.iters - io.iters) <= 2
    assert np.linalg.norm(x4 - xs) <= 1e-10 * np.linalg.norm(xs)


@pytest.mark.parametrize("jacobi", [False, True])
def test_fcg_trajectory_equals_pcg(jacobi):
    """Flexible CG with a fixed SPD preconditioner produces the iterates of
    preconditioned CG in exact arithmetic (its extra orthogonalisation term
    vanishes): the oracle's FCG and CG agree iterate by iterate, to rounding,
    for the identity and the Jacobi preconditioner."""
    A, S = lap2d(16)
    b = P.uniform(256, 7)
    s = O.Settings(maxiter=40, rtol=0.0, use_zero_guess=True)
    Pc = O.prec(1, A.dinv()) if jacobi else None
    xs_cg, res_cg, xs_fcg, res_fcg = [], [], [], []

    def rec(xs, rs):
        def f(x, res):
            xs.append(np.array(x))
            rs.append(res)
            return False
        return f

    O.cg(A, b, np.zeros(256), s, P=Pc, diag=rec(xs_cg, res_cg))
    O.fcg(A, b, np.zeros(256), s, P=Pc, diag=rec(xs_fcg, res_fcg))
    assert len(xs_cg) == len(xs_fcg) == 40
    for k in range(40):
        scale = np.linalg.norm(xs_cg[k])
        assert np.linalg.norm(xs_fcg[k] - xs_cg[k]) <= 1e-8 * scale, k
        assert abs(res_fcg[k] - res_cg[k]) <= 1e-6 * res_cg[0], k


def test_bicgstab_early_exit_tolerance_is_absolute():
    """bicgstab.hh:136-142 exits when |s| < rtol with rtol absolute (not
    rtol |b|) and reports iters = iter.  Pinned where the two tolerances
    disagree, on both sides: b scaled down (the absolute test fires first:
    early exit with a true residual below rtol but far above rtol |b|) and
    b scaled up (rtol <= |s| < rtol |b| at the converging iteration: the
    absolute test does not fire, the full step runs and iters = iter + 1).
    The planted relative-tolerance mistake changes the outcome in both."""
    n = 12
    rp, ci, v = P.convdiff3d(n, (10.0, 5.0, 2.5))
    A = O.Csr(rp, ci, v)
    S = sp.csr_matrix((v, ci, rp))
    b0 = P.uniform(n ** 3, 4)
    s = O.Settings(maxiter=500, rtol=1e-8, use_zero_guess=True)
    rtol = float(np.float32(1e-8))
    # b scaled down: |b| = 2.4e-4, so rtol |b| = 2.4e-12
    b = b0 * 1e-5
    x, info, h = O.bicgstab(A, b, np.zeros(n ** 3), s, hist_cap=500)
    tr = np.linalg.norm(b - S @ x)
    assert info.reason == "converged_rtol"
    assert tr < rtol and tr > 100 * rtol * np.linalg.norm(b)
    # the early exit skips the residual update: the last recorded |res| is
    # that of the previous iteration, still above rtol
    assert info.res_norm_final > rtol
    assert np.count_nonzero(h) == info.iters
    with O.mutant_relative_early_exit():
        xm, im, _ = O.bicgstab(A, b, np.zeros(n ** 3), s)
    assert im.iters > info.iters + 2
    # b scaled up: |b| = 2.4e6; the converging iteration's |s| lies between
    # rtol and rtol |b|, so only the planted mistake takes the early exit
    b = b0 * 1e5
    x, info, h = O.bicgstab(A, b, np.zeros(n ** 3), s, hist_cap=500)
    tol = rtol * np.linalg.norm(b)
    tr = np.linalg.norm(b - S @ x)
    assert info.reason == "converged_rtol"
    assert tr < 1.001 * tol
    assert h[info.iters - 1] < tol <= h[info.iters - 2]
    with O.mutant_relative_early_exit():
        xm, im, _ = O.bicgstab(A, b, np.zeros(n ** 3), s)
    assert im.iters == info.iters - 1
    assert not np.array_equal(xm, x)
