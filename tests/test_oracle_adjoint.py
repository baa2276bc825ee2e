"""Pins the oracle's backward-Euler adjoint (CPU only).

The reference has no adjoint, so the checks come from the mathematics:
  - the sampled outer product equals the dense outer product masked by the
    pattern (brute force);
  - the coefficient chain (dL/dA values -> dL/dlog k) equals central finite
    differences of the linear functional J(theta) = sum_e G_e A_e(exp theta)
    assembled by the problem generator, including single boundary, edge and
    corner nodes (corners touch no face of A: gradient exactly 0);
  - the full adjoint gradient equals central finite differences of the
    forward loss along seeded directions, for log k and for u^0.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from flecsolve_amd import problems as P
from oracle import oracle as O


def test_sampled_outer_brute_force():
    rng = np.random.default_rng(3)
    M = sp.random(40, 30, density=0.15, random_state=4, format="csr")
    M.sort_indices()
    lam, x = rng.standard_normal(40), rng.standard_normal(30)
    out = rng.standard_normal(M.nnz)
    ref = out.copy()
    O.sampled_outer(M.indptr, M.indices, lam, x, -0.75, out)
    dense = (-0.75 * lam)[:, None] * x[None, :]
    rows = np.repeat(np.arange(40), np.diff(M.indptr))
    np.testing.assert_array_equal(out, ref + dense[rows, M.indices])


def _linear_functional(n, theta, G, scale, shift=1.0):
    rp, ci, v = P.diffusion3d(n, np.exp(theta), scale=scale, shift=shift)
    return float(np.dot(G, v)), rp, ci


@pytest.mark.parametrize("n", [1, 3, 5])
def test_coeff_grad_matches_fd_of_linear_functional(n):
    rng = np.random.default_rng(n)
    NL = (n + 2) ** 3
    theta = 0.5 * rng.standard_normal(NL)
    scale = 2.7
    _, rp, ci = _linear_functional(n, theta, np.zeros(7 * n ** 3 - 6 * n * n), scale)
    G = rng.standard_normal(ci.size)
    g = O.diffusion3d_coeff_grad(n, rp, ci, G, np.exp(theta), scale)
    eps = 1e-5
    for trial in range(4):
        d = rng.uniform(-1, 1, NL)
        jp = _linear_functional(n, theta + eps * d, G, scale)[0]
        jm = _linear_functional(n, theta - eps * d, G, scale)[0]
        fd = (jp - jm) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-8 * max(1.0, abs(fd)), (trial, fd, g @ d)
    # single nodes: interior, face, edge, corner of the lattice
    L = n + 2
    for (i, j, k) in [(1, 1, 1), (0, 1, 1), (0, 0, 1), (0, 0, 0), (L - 1, L - 1, L - 1),
                      (n // 2 + 1, 0, n // 2 + 1)]:
        p = i + L * (j + L * k)
        e = np.zeros(NL)
        e[p] = 1.0
        fd = (_linear_functional(n, theta + eps * e, G, scale)[0] -
              _linear_functional(n, theta - eps * e, G, scale)[0]) / (2 * eps)
        assert abs(fd - g[p]) <= 1e-8 * max(1.0, abs(fd)), ((i, j, k), fd, g[p])
    corner = 0
    assert g[corner] == 0.0


def _heat(n, steps, seed=11, gamma=3.0):
    theta = np.log(P.lognormal_field(n, 1.0, 2.0, seed))
    theta_t = np.log(P.lognormal_field(n, 1.0, 2.0, seed + 1))
    u0 = P.bumps(n, 3, seed + 2)
    s = O.Settings(maxiter=2000, rtol=1e-14, use_zero_guess=False)

    def A_of(th):
        return O.Csr(*P.diffusion3d(n, np.exp(th), scale=gamma, shift=1.0))

    ut = O.be_forward(A_of(theta_t), u0, steps, s)[0][-1]

    def loss(th, u=u0):
        d = O.be_forward(A_of(th), u, steps, s)[0][-1] - ut
        return float(d @ d)

    return theta, u0, ut, s, A_of, loss, gamma


@pytest.mark.parametrize("n,steps", [(4, 3), (6, 5)])
def test_adjoint_gradient_matches_fd(n, steps):
    theta, u0, ut, s, A_of, loss, gamma = _heat(n, steps)
    A = A_of(theta)
    L0, G, gu0, uT = O.be_adjoint(A, u0, ut, steps, s)
    assert abs(L0 - loss(theta)) <= 1e-12 * L0
    g = O.diffusion3d_coeff_grad(n, A.rowptr, A.colind, G, np.exp(theta), gamma)
    rng = np.random.default_rng(100 + n)
    eps = 1e-4
    for trial in range(3):
        d = rng.uniform(-1, 1, theta.size)
        fd = (loss(theta + eps * d) - loss(theta - eps * d)) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-6 * abs(fd), (trial, fd, g @ d)
        du = rng.uniform(-1, 1, u0.size)
        fdu = (loss(theta, u0 + eps * du) - loss(theta, u0 - eps * du)) / (2 * eps)
        assert abs(fdu - gu0 @ du) <= 1e-6 * abs(fdu), (trial, fdu, gu0 @ du)


def test_planted_sign_error_is_caught():
    # the FD check above fails for a plausible mistake: dropping the minus
    # sign of the sampled outer product flips the coefficient gradient
    n, steps = 4, 3
    theta, u0, ut, s, A_of, loss, gamma = _heat(n, steps)
    A = A_of(theta)
    _, G, _, _ = O.be_adjoint(A, u0, ut, steps, s)
    g_bad = O.diffusion3d_coeff_grad(n, A.rowptr, A.colind, -G, np.exp(theta), gamma)
    d = np.random.default_rng(5).uniform(-1, 1, theta.size)
    eps = 1e-4
    fd = (loss(theta + eps * d) - loss(theta - eps * d)) / (2 * eps)
    assert abs(fd - g_bad @ d) > 1e-3 * abs(fd)
