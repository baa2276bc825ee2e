"""Helpers shared by the time-integrator tests: the oracle runs and the
parsing of tests/cpp/bdf_check and the heat example's output lines."""
from pathlib import Path

import numpy as np

from oracle import bdf as B
from oracle import oracle as O
from oracle import rk as R

GOLDEN = Path(__file__).resolve().parent / "golden"


def oracle_rate(lam, ic, n):
    """{section: (steps, rejects, time, max(u) as hex)} of the reference's
    rate test on the oracle."""
    secs = B.read_sections(GOLDEN / "bdf_rate.cfg")
    out = {}
    for name in ("bdf-2", "bdf-5"):
        op = B.RateOp(lam)
        ti = B.Integrator(B.settings_from_section(secs[name]), op, B.rate_solver(op), n)
        x = B.run(ti, np.full(n, ic))
        out[name] = (ti.get_current_step(), ti.num_step_rejects(),
                     float(ti.get_current_time()), float(x.max()).hex())
    return out


def parse_rate(text):
    out = {}
    for line in text.splitlines():
        t = line.split()
        if t and t[0] in ("bdf-2", "bdf-5"):
            out[t[0]] = (int(t[2]), int(t[4]), float(t[6]), float.fromhex(t[8]).hex())
    return out


def oracle_heat(n, alpha, cfg=GOLDEN / "heat_implicit.cfg"):
    """((steps, rejects, solves, iters), u) of the heat example on the oracle
    (stage solves on I - gamma F by the configured CG or BiCGSTAB)."""
    secs = B.read_sections(cfg)
    kind = secs["linear-solver"]["type"]
    lso = secs["linear-solver.options"]
    st = O.Settings(maxiter=int(lso["maxiter"]), rtol=float(lso["rtol"]),
                    use_zero_guess=lso["use-zero-guess"] == "true")
    rp, ci, fv, u0 = B.heat2d(n, alpha)
    op = B.CsrRhsOp(rp, ci, fv)
    solve = B.cg_stage_solver(op, st, kind)
    ti = B.Integrator(B.settings_from_section(secs["time-integrator"]), op, solve, u0.size)
    u = B.run(ti, u0)
    return (ti.get_current_step(), ti.num_step_rejects(), solve.solves, solve.iters), u


def parse_heat(text):
    """(steps, rejects, solves, iters) from a 'heat ...' summary line."""
    for line in text.splitlines():
        t = line.split()
        if t and t[0] == "heat":
            kv = dict(zip(t[1::2], t[2::2]))
            return (int(kv["steps"]), int(kv["rejects"]), int(kv["solves"]), int(kv["iters"]))
    raise AssertionError("no heat summary line:\n" + text)


def oracle_explicit(lam, ic, n):
    """{name: (steps, rejects, time, max(u) as hex)} of the reference's
    explicit test (RK23 / RK45, fixed and variable sections) on the oracle."""
    secs = B.read_sections(GOLDEN / "rk_rate.cfg")
    out = {}
    for cls, tag in ((R.RK23, "rk23"), (R.RK45, "rk45")):
        for sec in ("fixed", "variable"):
            ti = cls(R.settings_from_section(secs[sec]), lambda x: lam * x, n)
            x = R.run(ti, np.full(n, ic))
            out[f"{tag}-{sec}"] = (ti.get_current_step(), ti.total_step_rejects,
                                   float(ti.get_current_time()), float(x.max()).hex())
    return out


def parse_explicit(text):
    out = {}
    for line in text.splitlines():
        t = line.split()
        if t and t[0].startswith("rk"):
            out[t[0]] = (int(t[2]), int(t[4]), float(t[6]), float.fromhex(t[8]).hex())
    return out


def oracle_heat_explicit(n, alpha, cfg=GOLDEN / "heat_explicit.cfg"):
    """((steps, rejects), u) of the reference's explicit heat example (RK23,
    f = F u with the oracle's row-ordered spmv)."""
    secs = B.read_sections(cfg)
    rp, ci, fv, u0 = B.heat2d(n, alpha)
    F = O.Csr(rp, ci, fv)
    ti = R.RK23(R.settings_from_section(secs["time-integrator"]), F.spmv, u0.size)
    u = R.run(ti, u0)
    return (ti.get_current_step(), ti.total_step_rejects), u
