"""Backward-Euler heat problem u_t = div(k grad u) on the unit cube and the
gradient of a trajectory loss with respect to log k.

The time step is the reference's BE integrator path (time-integrators/
bdf.hh:74-95 with operator_adapter.hh:29-35; examples/heat_equation uses it
with a CG linear solver): A = I + (dt/h^2) L_k with h = 1/(n+1), Dirichlet
zero boundary, the 7-point operator with arithmetic-mean face coefficients.
Each step is a CG + Jacobi solve from the previous state.  The loss is
|u_T - u_target|^2, u_target the trajectory end of a second coefficient
field; its gradient comes from the library's discrete adjoint
(fsx_be_adjoint) chained to log k (fsx_diffusion3d_coeff_grad).  With
refine=True every solve is a mixed-precision solve on the symmetrically
scaled FP32 copy of A to the same tolerance: the restarted refinement
(method="ir": fsx_be_forward_ir / fsx_be_adjoint_ir, inner tolerance 1e-4)
or the FP32 CG with FP64 residual replacement (method="rr":
fsx_be_forward_rr / fsx_be_adjoint_rr, segment reduction 3e-2).  Every
solve, assembly and gradient kernel runs on the GPU through the C ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import fsx
from . import problems as P


class HeatProblem:
    def __init__(self, ctx: fsx.Context, n: int = 128, dt: float = 1e-3, steps: int = 50,
                 ckpt: int = 10, sigma: float = 1.0, ell: float = 8.0, bumps: int = 8,
                 rtol: float = 1e-12, maxiter: int = 5000, seeds: dict | None = None,
                 refine: bool = False, inner_rtol: float | None = None,
                 method: str = "ir"):
        s = seeds or P.seeds()
        if method not in ("ir", "rr"):
            raise ValueError("method must be 'ir' or 'rr'")
        self.refine, self.method = refine, method
        if inner_rtol is None:
            inner_rtol = 1e-4 if method == "ir" else 3e-2
        self.inner = fsx.Settings(maxiter=maxiter, rtol=inner_rtol)
        self.ctx, self.n, self.steps, self.ckpt = ctx, n, steps, ckpt
        self.gamma = dt * (n + 1) ** 2          # dt / h^2
        self.N = n ** 3
        self.NL = (n + 2) ** 3
        self.theta0 = np.log(P.lognormal_field(n, sigma, ell, s["heat_k"]))
        k_target = P.lognormal_field(n, sigma, ell, s["heat_k_target"])
        rp, ci, v = P.diffusion3d(n, k_target, scale=self.gamma, shift=1.0)
        self.A = ctx.csr(rp, ci, v)
        self.kvec = ctx.zeros(self.NL)
        self.dinv = ctx.zeros(self.N)
        self.u0 = ctx.vec(P.bumps(n, bumps, s["heat_u0"]))
        self.settings = fsx.Settings(maxiter=maxiter, rtol=rtol, use_zero_guess=False)
        self.A32 = self.A.convert(fsx.F32) if refine else None
        self._set_k(k_target)
        self.u_target, self.target_stats = self._forward()
        self._d = ctx.zeros(self.N)

    def _set_k(self, k: np.ndarray):
        self.kvec.upload(k)
        self.A.set_diffusion3d_values(self.n, self.kvec, 1.0, self.gamma)
        fsx.check(fsx.lib().fsx_csr_diag_inv(self.ctx.h, self.A.h, self.dinv.h), "diag_inv")
        if self.refine:
            self.A32.set_values_sym_scaled(self.A)

    def _forward(self):
        if self.refine:
            fwd = fsx.be_forward_rr if self.method == "rr" else fsx.be_forward_ir
            return fwd(self.A, self.A32, self.u0, self.steps, self.settings, self.inner)
        return fsx.be_forward(self.A, self.u0, self.steps, self.settings, self.dinv)

    def loss(self, theta: np.ndarray) -> float:
        """|u_T(theta) - u_target|^2 with k = exp(theta)."""
        self._set_k(np.exp(theta))
        uT, _ = self._forward()
        fsx.check(fsx.lib().fsx_axpy(self.ctx.h, -1.0, self.u_target.h, uT.h, self._d.h),
                  "axpy")
        return fsx.dot(self._d, self._d)

    def loss_and_grad(self, theta: np.ndarray):
        """(loss, dL/dtheta on the (n+2)^3 lattice, dL/du0, adjoint stats)."""
        self._set_k(np.exp(theta))
        if self.refine:
            adj = fsx.be_adjoint_rr if self.method == "rr" else fsx.be_adjoint_ir
            loss, G, uT, gu0, st = adj(self.A, self.A32, self.u0, self.u_target, self.steps,
                                       self.ckpt, self.settings, self.inner)
        else:
            loss, G, uT, gu0, st = fsx.be_adjoint(self.A, self.u0, self.u_target, self.steps,
                                                  self.ckpt, self.settings, self.dinv)
        g = self.A.diffusion3d_coeff_grad(self.n, G, self.kvec, self.gamma)
        return loss, g.get(), gu0.get(), st
