"""The C++ header API (include/flecsolve) used the way the reference's tests
use theirs: tests/cpp/cg_dropin.cpp mirrors flecsolve/solvers/test/cg.cc with
its A-norm / Trefethen-Bau diagnostic, plus GMRES(30)+Dinv, BiCGSTAB, CG with
a hybrid Gauss-Seidel preconditioner, CG / GMRES / BiCGSTAB chosen by
krylov_factory from config files, CG, the mixed-precision refinement and
the FP32 CG with FP64 residual replacement (ir::rr) through the fused solver
policies, CG over the parallel CSR operator as
flecsolve/matrices/test/par_csr.cc runs it (b = 0, x0 = 2; 1 and 4 colours),
and CG on a Matrix Market file read through the matrix_market / coo headers
(gpu)."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "cpp" / "cg_dropin"


def test_cpp_dropin_solvers():
    assert BIN.exists(), "build first: python -m flecsolve_amd.build"
    gold = Path(__file__).resolve().parent / "golden" / "sym5.mtx"
    r = subprocess.run([str(BIN), str(gold), str(gold.parent)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
