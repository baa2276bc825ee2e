"""The C++ time-integrator headers (include/flecsolve/time-integrators)
over host vectors against the oracle (no device use): tests/cpp/bdf_check_host
runs the reference's rate test and its explicit RK23 / RK45 test, and the
rate problem under every method, predictor, controller, step strategy and
error scaling (bit-identical to the oracle: same steps, rejections and final
value) and the heat example with krylov_factory CG
stage solves (same steps, rejections, solves and CG iterations; solution
within 1e-12), and rejects malformed integrator configs."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import bdf_common as BC

MATRIX = BC.integrator_matrix()

HOST = Path(__file__).resolve().parent / "cpp" / "bdf_check_host"


def run(args, ok=True):
    assert HOST.exists(), "build first: python -m flecsolve_amd.build"
    r = subprocess.run([str(HOST)] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=120)
    if ok:
        assert r.returncode == 0, r.stdout + r.stderr
    return r


def test_rate_bit_identical():
    out = run(["rate", BC.GOLDEN / "bdf_rate.cfg", -1.0, 3.0, 5]).stdout
    assert BC.parse_rate(out) == BC.oracle_rate(-1.0, 3.0, 5)


def test_explicit_rk_bit_identical():
    out = run(["explicit", BC.GOLDEN / "rk_rate.cfg", -1.0, 3.0, 1]).stdout
    assert BC.parse_explicit(out) == BC.oracle_explicit(-1.0, 3.0, 1)


def test_heat_generic_cg(tmp_path):
    n = 40
    f = tmp_path / "u.bin"
    out = run(["heat", BC.GOLDEN / "heat_implicit.cfg", n, 1.0, f]).stdout
    got = BC.parse_heat(out)
    ref, u = BC.oracle_heat(n, 1.0)
    assert got == ref
    uh = np.fromfile(f)
    assert np.abs(uh - u).max() <= 1e-12 * np.abs(u).max()


@pytest.mark.parametrize("drop,msg", [("integrator", "required option"),
                                      ("predictor", "required option")])
def test_config_errors(tmp_path, drop, msg):
    lines = [l for l in (BC.GOLDEN / "bdf_rate.cfg").read_text().splitlines()
             if not l.startswith(drop + " ")]
    cfg = tmp_path / "bad.cfg"
    cfg.write_text("\n".join(lines) + "\n")
    r = run(["rate", cfg, -1.0, 3.0, 2], ok=False)
    assert r.returncode == 1 and msg in r.stderr, r.stderr


def test_config_bad_values(tmp_path):
    text = (BC.GOLDEN / "bdf_rate.cfg").read_text()
    for old, new, msg in (("integrator = BDF5", "integrator = BDF7", "invalid value"),
                          ("starting-integrator = CN", "starting-integrator = BDF2",
                           "starting integrator"),
                          ("problem-scales = 1.", "problem-scales = 1. x", "invalid value")):
        assert old in text
        cfg = tmp_path / "bad.cfg"
        cfg.write_text(text.replace(old, new, 1))
        r = run(["rate", cfg, -1.0, 3.0, 2], ok=False)
        assert r.returncode == 1 and msg in r.stderr, (new, r.stderr)


@pytest.mark.parametrize("name", sorted(MATRIX))
def test_integrator_branches_bit_identical(tmp_path, name):
    cfg = tmp_path / "one.cfg"
    BC.write_section(cfg, name, MATRIX[name])
    out = run(["rateone", cfg, name, -1.0, 3.0, 4]).stdout
    assert BC.parse_rate_one(out, name) == BC.oracle_rate_one(name, MATRIX[name], -1.0, 3.0, 4)
