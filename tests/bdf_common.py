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


# Integrator configurations beyond the reference's own test: every method,
# predictor, controller, step strategy and error scaling (rate problem
# u' = -u, u(0) = 3, to t = 1).
_BASE = {"initial-time": "0.0", "final-time": "1.0", "initial-dt": "0.1",
         "max-steps": "1000", "min-dt": "0.001", "max-dt": "0.5",
         "trunc-error-rtol": "1e-6", "trunc-error-atol": "1e-8",
         "time-trunc-error-norm": "inf", "time-error-scaling": "fixed-resolution"}


def integrator_matrix():
    """name -> {key: value} (one config section each)."""
    m = {}
    for meth, start in (("BE", "BE"), ("BDF2", "BE"), ("BDF3", "CN"), ("BDF4", "CN")):
        for ctl in ("PC.4.7", "H211b", "PC11", "Deadbeat"):
            m[f"{meth}-{start}-{ctl}"] = dict(
                _BASE, **{"integrator": meth, "starting-integrator": start,
                          "predictor": "leapfrog",
                          "timestep-selection-strategy": "truncation-error",
                          "pi-controller-type": ctl})
    # CN (AB2 predictor) and BDF6 only at constant steps: under truncation-error
    # control the reference's CN run shrinks dt to ~2e-6 (its AB2 predictor
    # reads a stored f(u_n) that is refreshed only on the first step) and its
    # BDF6 (sixth coefficient as written) until two step sizes coincide
    m["CN-ab2-constant"] = dict(_BASE, **{"integrator": "CN", "starting-integrator": "CN",
                                          "predictor": "ab2", "initial-dt": "0.0625",
                                          "timestep-selection-strategy": "constant"})
    m["BDF6-constant"] = dict(_BASE, **{"integrator": "BDF6", "starting-integrator": "CN",
                                        "predictor": "leapfrog", "initial-dt": "0.0625",
                                        "timestep-selection-strategy": "constant"})
    m["BDF3-deadbeat-nopi"] = dict(_BASE, **{
        "integrator": "BDF3", "starting-integrator": "CN", "predictor": "leapfrog",
        "timestep-selection-strategy": "truncation-error", "use-pi-controller": "false"})
    m["BDF2-l2-fixed-scaling"] = dict(_BASE, **{
        "integrator": "BDF2", "starting-integrator": "CN", "predictor": "leapfrog",
        "timestep-selection-strategy": "truncation-error", "time-trunc-error-norm": "l2",
        "time-error-scaling": "fixed-scaling", "problem-scales": "1.0",
        "use-initial-predictor": "false"})
    m["BDF4-constant"] = dict(_BASE, **{
        "integrator": "BDF4", "starting-integrator": "CN", "predictor": "leapfrog",
        "timestep-selection-strategy": "constant", "initial-dt": "0.0625"})
    m["BDF2-limit-change"] = dict(_BASE, **{
        "integrator": "BDF2", "starting-integrator": "BE", "predictor": "leapfrog",
        "timestep-selection-strategy": "limit-relative-change",
        "target-relative-change": "0.05"})
    m["BDF2-final-constant"] = dict(_BASE, **{
        "integrator": "BDF2", "starting-integrator": "CN", "predictor": "leapfrog",
        "timestep-selection-strategy": "final-constant", "number-of-time-intervals": "5",
        "initial-dt": "0.05", "max-dt": "0.2"})
    return m


def write_section(path, name, kv):
    path.write_text(f"[{name}]\n" + "".join(f"{k} = {v}\n" for k, v in kv.items()))


def oracle_rate_one(name, kv, lam, ic, n):
    """(steps, rejects, time, max(u) hex) of one configuration on the oracle,
    with the final-constant counter starting afresh (one process per run on
    the C++ side)."""
    B.Integrator._final_constant_i = 1
    op = B.RateOp(lam)
    ti = B.Integrator(B.settings_from_section(kv), op, B.rate_solver(op), n)
    x = B.run(ti, np.full(n, ic))
    return (ti.get_current_step(), ti.num_step_rejects(), float(ti.get_current_time()),
            float(x.max()).hex())


def parse_rate_one(text, name):
    for line in text.splitlines():
        t = line.split()
        if t and t[0] == name:
            return (int(t[2]), int(t[4]), float(t[6]), float.fromhex(t[8]).hex())
    raise AssertionError("no line for " + name + ":\n" + text)
