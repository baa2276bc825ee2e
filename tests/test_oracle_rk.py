"""Oracle RK23 / RK45 (oracle/rk.py) against the reference's own
known-answer test and what the mathematics fixes.

Pins:
  * time-integrators/test/explicit.cc:80-101 with the reference's
    explicit.cfg (tests/golden/rk_rate.cfg): u' = -u, u(0) = 3 on a one-cell
    mesh to t = 1 -- fixed steps: RK23 within 1e-5, RK45 within 1e-9;
    variable steps: RK23 in exactly 32 steps within 1e-5, RK45 in exactly 6
    steps within 1e-6;
  * global order 3 (RK23) and 5 (RK45, which advances with its fifth-order
    solution) at fixed steps;
  * the embedded error estimate of the last step shrinks as dt^3 and dt^5.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import bdf as B
from oracle import rk as R

GOLDEN = Path(__file__).resolve().parent / "golden"


def rate_run(cls, section):
    secs = B.read_sections(GOLDEN / "rk_rate.cfg")
    ti = cls(R.settings_from_section(secs[section]), lambda x: -1.0 * x, 1)
    x = R.run(ti, np.full(1, 3.0))
    return ti, abs(3.0 * math.exp(-ti.get_final_time()) - x.max())


def test_reference_known_answer():
    ti, err = rate_run(R.RK23, "fixed")
    assert ti.get_final_time() == 1.0 and err < 1e-5
    ti, err = rate_run(R.RK45, "fixed")
    assert ti.get_final_time() == 1.0 and err < 1e-9
    ti, err = rate_run(R.RK23, "variable")
    assert ti.get_final_time() == 1.0 and err < 1e-5 and ti.get_current_step() == 32
    ti, err = rate_run(R.RK45, "variable")
    assert ti.get_final_time() == 1.0 and err < 1e-6 and ti.get_current_step() == 6


def fixed(cls, dt):
    s = R.Settings(final_time=1.0, initial_dt=dt, max_dt=dt, use_fixed_dt=True,
                   max_steps=10 ** 6)
    ti = cls(s, lambda x: -1.0 * x, 1)
    x = R.run(ti, np.ones(1))
    assert ti.get_current_time() == 1.0
    return abs(x[0] - math.exp(-1.0)), float(np.abs(ti.z).max())


@pytest.mark.parametrize("cls,order", [(R.RK23, 3), (R.RK45, 5)])
def test_orders_and_estimates(cls, order):
    e1, z1 = fixed(cls, 1 / 16)
    e2, z2 = fixed(cls, 1 / 32)
    assert abs(math.log2(e1 / e2) - order) < 0.2, (e1, e2)
    assert abs(math.log2(z1 / z2) - order) < 0.2, (z1, z2)


def test_float_settings():
    s = R.settings_from_section({"initial-time": "0", "final-time": "1", "max-steps": "5",
                                 "atol": "1e-5", "safety-factor": "0.9"})
    assert s.atol == float(np.float32(1e-5)) != 1e-5
    assert s.safety_factor == float(np.float32(0.9))
