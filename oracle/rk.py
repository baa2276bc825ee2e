"""CPU restatement of the reference's explicit Runge-Kutta integrators.

TEST INFRASTRUCTURE ONLY: imported by tests/ and nowhere else.

Follows flecsolve/time-integrators/rk23.hh (Bogacki-Shampine 2(3),
rk23.hh:56-142) and rk45.hh (Runge-Kutta-Fehlberg 4(5), rk45.hh:42-147)
operation for operation on numpy float64 vectors (each vector operation one
elementwise expression in the reference's operand order), with the float
settings of the reference (safety-factor and atol are float, rk23.hh:25-29).

Pins (tests/test_oracle_rk.py): the reference's own known-answer test
(time-integrators/test/explicit.cc:80-101 with explicit.cfg): u' = -u,
u(0) = 3 on a one-cell mesh to t = 1 -- fixed-step RK23 within 1e-5 and RK45
within 1e-9, variable-step RK23 in exactly 32 steps within 1e-5 and RK45 in
exactly 6 steps within 1e-6; global order 3 and 5 at fixed steps; the
embedded estimates shrink as dt^3 and dt^5.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Settings:
    """rk23::settings / rk45::settings with the option defaults
    (rk23.hh:35-47; base keys parameters.hh:41-55)."""
    initial_time: float = 0.0
    final_time: float = 1.0
    max_steps: int = 1000
    max_dt: float = np.finfo(np.float64).max
    min_dt: float = np.finfo(np.float64).tiny
    initial_dt: float = 0.0
    safety_factor: float = float(np.float32(0.9))
    atol: float = float(np.float32(1e-9))
    use_fixed_dt: bool = False


def settings_from_section(kv: dict) -> Settings:
    s = Settings()
    conv = {"initial-time": ("initial_time", float), "final-time": ("final_time", float),
            "max-steps": ("max_steps", int), "max-dt": ("max_dt", float),
            "min-dt": ("min_dt", float), "initial-dt": ("initial_dt", float),
            "safety-factor": ("safety_factor", lambda t: float(np.float32(float(t)))),
            "atol": ("atol", lambda t: float(np.float32(float(t)))),
            "use-fixed-dt": ("use_fixed_dt",
                             lambda t: t.strip().lower() in ("true", "1", "yes", "on"))}
    for k, text in kv.items():
        name, f = conv[k]
        setattr(s, name, f(text))
    return s


class _Base:
    """base.hh:24-60: time, dt and accepted-step bookkeeping."""

    def __init__(self, s: Settings, f, n: int, nwork: int):
        self.p, self.f = s, f
        self.current_time = s.initial_time
        self.current_dt = s.initial_dt
        self.integrator_step = 0
        self.total_step_rejects = 0
        self.w = [np.zeros(n) for _ in range(nwork)]

    def get_current_time(self):
        return self.current_time

    def get_final_time(self):
        return self.p.final_time

    def get_current_dt(self):
        return self.current_dt

    def get_current_step(self):
        return self.integrator_step

    def fixed_dt(self):
        return self.p.use_fixed_dt

    def update(self):
        self.current_time += self.current_dt
        self.integrator_step += 1

    def _assert_can_advance(self):
        assert self.integrator_step < self.p.max_steps and \
            self.current_time < self.p.final_time, "already finished integrating"

    def check_solution(self):
        err = float(np.sqrt(np.dot(self.z, self.z)))
        return err < self.p.atol or abs(self.current_dt - self.p.min_dt) < 1e-10


class RK23(_Base):
    def __init__(self, s, f, n):
        super().__init__(s, f, n, 6)

    @property
    def z(self):
        return self.w[4]

    def advance(self, dt, curr, out):
        """rk23.hh:71-103"""
        self._assert_can_advance()
        self.current_dt = dt
        F = self.f
        k1, k2, k3, k4, z, nx = self.w
        k1[:] = F(curr)
        nx[:] = (0.5 * dt) * k1 + curr
        k2[:] = F(nx)
        nx[:] = (0.75 * dt) * k2 + curr
        k3[:] = F(nx)
        nx[:] = 2.0 * k1 + 3.0 * k2
        nx[:] = 4.0 * k3 + nx
        nx[:] = (dt / 9.0) * nx + curr
        k4[:] = F(nx)
        z[:] = -5.0 * k1 + 6.0 * k2
        z[:] = 8.0 * k3 + z
        z[:] = -9.0 * k4 + z
        z[:] = (dt / 72.0) * z
        out[:] = nx

    def get_next_dt(self, good):
        """rk23.hh:116-136"""
        p = self.p
        if p.use_fixed_dt:
            return min(self.current_dt, p.final_time - self.current_time)
        err = float(np.sqrt(np.dot(self.z, self.z)))
        nd = p.safety_factor * self.current_dt * (p.atol / err) ** (1.0 / 3.0)
        nd = min(max(nd, p.min_dt), p.max_dt)
        nd = min(nd, p.final_time - self.current_time)
        if not good:
            self.total_step_rejects += 1
        return nd


class RK45(_Base):
    def __init__(self, s, f, n):
        super().__init__(s, f, n, 8)

    @property
    def z(self):
        return self.w[6]

    def advance(self, dt, curr, out):
        """rk45.hh:56-111"""
        self._assert_can_advance()
        self.current_dt = dt
        F = self.f
        k1, k2, k3, k4, k5, k6, z, nx = self.w
        k1[:] = F(curr)
        nx[:] = (0.25 * dt) * k1 + curr
        k2[:] = F(nx)
        nx[:] = (3.0 * dt / 32.) * k1 + curr
        nx[:] = (9.0 * dt / 32.) * k2 + nx
        k3[:] = F(nx)
        nx[:] = (1932. * dt / 2197.) * k1 + curr
        nx[:] = (-7200. * dt / 2197.) * k2 + nx
        nx[:] = (7296. * dt / 2197.) * k3 + nx
        k4[:] = F(nx)
        nx[:] = (439. * dt / 216.) * k1 + curr
        nx[:] = (-8. * dt) * k2 + nx
        nx[:] = (3680. * dt / 513.) * k3 + nx
        nx[:] = (-845. * dt / 4104.) * k4 + nx
        k5[:] = F(nx)
        nx[:] = (-8. * dt / 27.) * k1 + curr
        nx[:] = (2. * dt) * k2 + nx
        nx[:] = (-3544. * dt / 2565.) * k3 + nx
        nx[:] = (1859. * dt / 4104.) * k4 + nx
        nx[:] = (-11. * dt / 40.) * k5 + nx
        k6[:] = F(nx)
        z[:] = (25. * dt / 216.) * k1 + curr
        z[:] = (1408. * dt / 2565.) * k3 + z
        z[:] = (2197. * dt / 4104.) * k4 + z
        z[:] = (-0.2 * dt) * k5 + z
        nx[:] = (16. * dt / 135.) * k1 + curr
        nx[:] = (6656. * dt / 12825.) * k3 + nx
        nx[:] = (28561. * dt / 56430.) * k4 + nx
        nx[:] = (-9. * dt / 50.) * k5 + nx
        nx[:] = (2. * dt / 55.) * k6 + nx
        z[:] = nx - z
        out[:] = nx

    def get_next_dt(self, good):
        """rk45.hh:124-142"""
        p = self.p
        if p.use_fixed_dt:
            return min(self.current_dt, p.final_time - self.current_time)
        if good:
            err = float(np.sqrt(np.dot(self.z, self.z)))
            nd = p.safety_factor * self.current_dt * (p.atol / err) ** (1.0 / 5.0)
            nd = min(max(nd, p.min_dt), p.max_dt)
            return min(nd, p.final_time - self.current_time)
        self.total_step_rejects += 1
        return p.safety_factor * self.current_dt


def run(ti, u0):
    """The time loop of time-integrators/test/explicit.cc:71-80."""
    x = np.array(u0, dtype=np.float64, copy=True)
    xnew = np.zeros_like(x)
    dt = ti.get_current_dt()
    while ti.get_current_time() < ti.get_final_time():
        ti.advance(dt, x, xnew)
        good = ti.check_solution()
        if good or ti.fixed_dt():
            ti.update()
            x, xnew = xnew, x
        dt = ti.get_next_dt(good)
    return x
