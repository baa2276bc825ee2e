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
  error cases;
* tests/cpp/amg_setup_check.cpp: the library's UA-AMG setup header on 7
  host threads, under ASan + UBSan and separately under ThreadSanitizer,
  its output identical to a single-thread run's.

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


def _setup_input(d):
    from flecsolve_amd import problems as P
    import numpy as np
    rp, ci, v = P.diffusion3d(28, P.lognormal_field(28, 1.0, 3.0, 21))
    d.mkdir()
    rp.astype(np.int64).tofile(d / "rp.bin")
    ci.astype(np.int64).tofile(d / "ci.bin")
    v.tofile(d / "v.bin")
    return d


@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_amg_setup_threads(tmp_path, san):
    """The UA-AMG setup header (flecsolve_amd/csrc/fsx_amg_setup.hpp, through
    tests/cpp/amg_setup_check.cpp) with 7 host threads under ASan + UBSan and
    under ThreadSanitizer (the row-block threads share the input read-only
    and write disjoint rows): no report, and every output file identical to
    the unsanitised single-thread run's."""
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
             "-pthread", f"-fsanitize={san}", "-fno-sanitize-recover=all",
             "-I", str(ROOT / "flecsolve_amd" / "csrc")]
    exe = tmp_path / "amg_setup_san"
    r = subprocess.run(["g++", *flags, str(CPP / "amg_setup_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    plain = tmp_path / "plain"
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-pthread",
                        "-I", str(ROOT / "flecsolve_amd" / "csrc"),
                        str(CPP / "amg_setup_check.cpp"), "-o", str(plain)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    a, b = _setup_input(tmp_path / "a"), _setup_input(tmp_path / "b")
    env = dict(ENV, FSX_SETUP_THREADS="7", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(exe), str(a), "0.25", "3", "8", "5"], capture_output=True,
                       text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "WARNING: ThreadSanitizer" not in out and "runtime error" not in out, out
    r = subprocess.run([str(plain), str(b), "0.25", "3", "8", "5"], capture_output=True,
                       text=True, timeout=300, env=dict(os.environ, FSX_SETUP_THREADS="1"))
    assert r.returncode == 0, r.stderr
    names = sorted(p.name for p in b.iterdir())
    assert names == sorted(p.name for p in a.iterdir()) and "aggt0.bin" in names
    for n in names:
        assert (a / n).read_bytes() == (b / n).read_bytes(), n
