"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no device
use).  Builds with g++ -fsanitize=address,undefined into a temporary
directory and runs:

* tests/cpp/sanitize_host.cpp: the oracle's solvers (CG, FCG, GMRES,
  BiCGSTAB with identity, Dinv and weighted-Jacobi preconditioners; CG over
  colours), Jacobi and hybrid Gauss-Seidel relaxation, the UA-AMG setup and
  cycles, COO -> CSR and the Matrix Market reader, and the library's host
  input path fsx_io.cpp (header, entries, tocsr, fsx_csr_from_mm) on the
  golden files and on malformed files;
* tests/cpp/bdf_check.cpp over host vectors: the reference's BDF rate test
  and the heat example with krylov_factory CG stage solves (time-integrator
  and factory headers);
* tests/cpp/config_check.cpp: the config reader on the golden files and its
  error cases.

Any sanitizer report makes the program exit non-zero (halt_on_error, and
UBSan without recovery)."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CPP = ROOT / "tests" / "cpp"
GOLDEN = ROOT / "tests" / "golden"
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-I", str(ROOT / "include"), "-I", str(CPP)]
ENV = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def build(tmp, out, *srcs, extra=()):
    exe = tmp / out
    r = subprocess.run(["g++", *FLAGS, *extra, *map(str, srcs), "-o", str(exe)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return exe


def run(exe, *args):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True,
                       timeout=300, env=ENV)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out, out
    return r.stdout


def test_oracle_and_host_input_path(tmp_path):
    exe = build(tmp_path, "sanitize_host", CPP / "sanitize_host.cpp",
                ROOT / "oracle" / "flecsolve_oracle.cpp",
                ROOT / "flecsolve_amd" / "csrc" / "fsx_io.cpp", extra=["-fopenmp"])
    assert "PASS" in run(exe, GOLDEN)


def test_time_integrator_and_factory_headers(tmp_path):
    exe = build(tmp_path, "bdf_check_host", CPP / "bdf_check.cpp",
                ROOT / "flecsolve_amd" / "csrc" / "problems.cpp",
                extra=["-DBDF_HOST", "-fopenmp"])
    out = run(exe, "rate", GOLDEN / "bdf_rate.cfg", -1.0, 3.0, 3)
    assert out.strip()
    out = run(exe, "heat", GOLDEN / "heat_implicit.cfg", 16, 1.0, tmp_path / "u.bin")
    assert out.strip()


def test_config_reader(tmp_path):
    exe = build(tmp_path, "config_check", CPP / "config_check.cpp")
    assert "PASS" in run(exe, GOLDEN)
