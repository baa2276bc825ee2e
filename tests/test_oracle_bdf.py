"""Oracle BDF / CN integrator (oracle/bdf.py) against the reference's own
known-answer test and what the mathematics fixes.

Pins:
  * time-integrators/test/implicit.cc:121-134 with the reference's
    implicit.cfg (tests/golden/bdf_rate.cfg): u' = -u, u(0) = 3 to t = 1;
    BDF2 within 1e-3 in <= 49 steps with <= 7 rejected, BDF5 within 1e-7 in
    exactly 48 steps with exactly 14 rejected;
  * variable-step BDF2-5 coefficients are exact for polynomials of degree
    <= k; BDF6's sixth coefficient is the reference's (h1 - h5 in its
    denominator): scaled by (h1 - h5) / (h1 - h6) the formula is exact;
  * global convergence order at constant steps: BE 1, CN 2, BDF2 2, BDF3 3;
  * the heat equation started on a Laplacian eigenvector (maximum 1) takes
    the steps and rejections of the scalar rate problem with its eigenvalue,
    and ends at the same value;
  * CG stage solves of the heat example converge and the solution decays
    monotonically in max norm (maximum principle).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import bdf as B
from oracle import oracle as O

from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden"


def rate_run(section_kv, lam=-1.0, ic=3.0, n=4):
    s = B.settings_from_section(section_kv)
    op = B.RateOp(lam)
    ti = B.Integrator(s, op, B.rate_solver(op), n)
    x = B.run(ti, np.full(n, ic))
    return ti, x


def test_reference_known_answer():
    secs = B.read_sections(GOLDEN / "bdf_rate.cfg")
    ti, x = rate_run(secs["bdf-2"])
    err = abs(3.0 * math.exp(-ti.get_final_time()) - x.max())
    assert ti.get_final_time() == 1.0 and err < 1e-3
    assert ti.get_current_step() <= 49 and ti.num_step_rejects() <= 7
    ti, x = rate_run(secs["bdf-5"])
    err = abs(3.0 * math.exp(-ti.get_final_time()) - x.max())
    assert ti.get_final_time() == 1.0 and err < 1e-7
    assert ti.get_current_step() == 48 and ti.num_step_rejects() == 14


@pytest.mark.parametrize("m", [B.BDF2, B.BDF3, B.BDF4, B.BDF5, B.BDF6])
def test_coefficients_polynomial_exactness(m):
    rng = np.random.default_rng(10 + m)
    k = B.memory_size(m)
    for _ in range(20):
        steps = rng.uniform(0.2, 1.5, k)  # t_{n+1} - t_n, t_n - t_{n-1}, ...
        h = list(np.cumsum(steps))
        a, a0, _ = B.coefficients(m, h)
        a = list(a)
        if m == B.BDF6:
            a[5] *= (h[0] - h[4]) / (h[0] - h[5])
        # p(t_{n+1}) + sum a_i p(t_{n+1} - h_i) = gamma p'(t_{n+1}), deg p <= k
        assert abs(1.0 + sum(a)) < 1e-12 * max(1.0, max(abs(c) for c in a))
        for deg in range(1, k + 1):
            lhs = sum(ai * (-hi) ** deg for ai, hi in zip(a, h))
            rhs = a0 if deg == 1 else 0.0
            scale = sum(abs(ai) * hi ** deg for ai, hi in zip(a, h))
            assert abs(lhs - rhs) <= 1e-11 * scale, (m, deg, lhs, rhs)


def test_bdf6_coefficient_differs_from_textbook():
    h = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    a, a0, _ = B.coefficients(B.BDF6, h)
    lhs = sum(ai * (-hi) for ai, hi in zip(a, h))
    assert abs(lhs - a0) > 1e-3 * a0  # not exact: the reference's h1 - h5


def constant_run(integrator, starting, predictor, dt, lam=-1.0):
    s = B.Settings(initial_time=0.0, final_time=1.0, max_steps=100000, initial_dt=dt,
                   max_dt=dt, timestep_strategy=B.CONSTANT, integrator=integrator,
                   starting_integrator=starting, predictor=predictor,
                   time_error_scaling=B.FIXED_RESOLUTION)
    op = B.RateOp(lam)
    ti = B.Integrator(s, op, B.rate_solver(op), 2)
    x = B.run(ti, np.ones(2))
    assert ti.get_current_time() == 1.0
    return abs(x[0] - math.exp(lam))


@pytest.mark.parametrize("integrator,starting,predictor,order", [
    (B.BE, B.BE, B.LEAPFROG, 1), (B.CN, B.CN, B.AB2, 2),
    (B.BDF2, B.CN, B.LEAPFROG, 2), (B.BDF3, B.CN, B.LEAPFROG, 3)])
def test_constant_step_orders(integrator, starting, predictor, order):
    e1 = constant_run(integrator, starting, predictor, 1 / 64)
    e2 = constant_run(integrator, starting, predictor, 1 / 128)
    assert abs(math.log2(e1 / e2) - order) < 0.3, (e1, e2)


def test_heat_eigenvector_reduces_to_rate():
    n = 15
    rp, ci, fv, _ = B.heat2d(n, 1.0)
    h = 10.0 / (n + 1)
    lam = -(1.0 / (h * h)) * (4.0 - 4.0 * math.cos(math.pi / (n + 1)))
    s1 = np.sin(math.pi * np.arange(1, n + 1) / (n + 1))
    v = np.outer(s1, s1).ravel()
    assert v.max() == 1.0
    Fm = sp.csr_matrix((fv, ci, rp))
    assert np.abs(Fm @ v - lam * v).max() < 1e-12 * abs(lam)
    kv = B.read_sections(GOLDEN / "heat_implicit.cfg")["time-integrator"]

    op = B.CsrRhsOp(rp, ci, fv)

    def exact(b, x):
        x[:] = spla.spsolve(sp.csr_matrix((op.stage().vals, ci, rp)).tocsc(), b)
        return True

    ti = B.Integrator(B.settings_from_section(kv), op, exact, n * n)
    u = B.run(ti, v)
    ro = B.RateOp(lam)
    tr = B.Integrator(B.settings_from_section(kv), ro, B.rate_solver(ro), 1)
    x = B.run(tr, np.ones(1))
    assert (ti.get_current_step(), ti.num_step_rejects()) == \
        (tr.get_current_step(), tr.num_step_rejects())
    assert abs(u.max() - x[0]) <= 1e-10 * abs(x[0])


def test_heat_cg_stage_solves():
    secs = B.read_sections(GOLDEN / "heat_implicit.cfg")
    lso = secs["linear-solver.options"]
    st = O.Settings(maxiter=int(lso["maxiter"]), rtol=float(lso["rtol"]),
                    use_zero_guess=lso["use-zero-guess"] == "true")
    rp, ci, fv, u0 = B.heat2d(24, 1.0)
    op = B.CsrRhsOp(rp, ci, fv)
    solve = B.cg_stage_solver(op, st)
    ti = B.Integrator(B.settings_from_section(secs["time-integrator"]), op, solve, u0.size)
    maxes = []
    u = B.run(ti, u0, on_step=lambda t, x: maxes.append(np.abs(x).max()))
    assert abs(ti.get_current_time() - 0.1) < 1e-15
    assert solve.solves == ti.get_current_step() + ti.num_step_rejects()
    assert all(b <= a * (1 + 1e-9) for a, b in zip([50.0] + maxes, maxes))
    assert 0 < u.max() < 50.0


def test_validate_rejects_bad_settings():
    with pytest.raises(AssertionError):
        B.Settings(starting_integrator=B.BDF2).validate()
    with pytest.raises(AssertionError):  # fixed scaling needs problem scales
        B.Settings(time_error_scaling=B.FIXED_SCALING).validate()
    with pytest.raises(AssertionError):  # BDF needs the leapfrog predictor
        B.Settings(predictor=B.AB2, problem_scales=[1.0]).validate()
