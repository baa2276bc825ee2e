"""The UA-AMG setup on the host alone against the oracle (cpu).

tests/cpp/amg_setup_check runs flecsolve_amd/csrc/fsx_amg_setup.hpp -- the
pairwise aggregation and Galerkin coarsening fsx_amg_create runs -- without
a device.  Bar: every level's fine-row -> aggregate map and coarse matrix
identical to the oracle's restatement of mg::ua::bound_solver's setup
(coarsen.hh:26-655; bit for bit), with 1 and with 7 host threads (the
per-row parts run on contiguous row blocks, so the thread count must not
change a bit), on a lognormal 3D diffusion matrix large enough for several
threads per level, double and triple pairwise, and on a 2D Laplacian.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from flecsolve_amd import problems as P
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent
CHECK = ROOT / "tests" / "cpp" / "amg_setup_check"


def hierarchy(tmp_path, rp, ci, v, passes, threads, max_levels=8, min_coarse=5):
    d = tmp_path / f"h{passes}_{threads}"
    d.mkdir()
    np.asarray(rp, np.int64).tofile(d / "rp.bin")
    np.asarray(ci, np.int64).tofile(d / "ci.bin")
    np.asarray(v, np.float64).tofile(d / "v.bin")
    env = {"FSX_SETUP_THREADS": str(threads), "PATH": "/usr/bin:/bin"}
    r = subprocess.run([str(CHECK), str(d), "0.25", str(passes), str(max_levels),
                        str(min_coarse)], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    nl = int((d / "levels.txt").read_text())
    aggt = [np.fromfile(d / f"aggt{l}.bin", np.int64) for l in range(nl - 1)]
    mats = [None] + [(np.fromfile(d / f"rp{l}.bin", np.int64), np.fromfile(d / f"ci{l}.bin", np.int64),
                      np.fromfile(d / f"v{l}.bin", np.float64)) for l in range(1, nl)]
    return nl, aggt, mats


@pytest.mark.parametrize("case,passes", [("d3d", 2), ("d3d", 3), ("p2d", 2)])
def test_host_setup_matches_oracle(tmp_path, case, passes):
    if case == "d3d":
        rp, ci, v = P.diffusion3d(28, P.lognormal_field(28, 1.0, 3.0, 21))
    else:
        rp, ci, v = P.poisson2d(160)
    rp, ci = rp.astype(np.int64), ci.astype(np.int64)
    Mo = O.Amg(O.Csr(rp, ci, v), O.AmgSettings(max_levels=8, maxiter=1, atol=0.0,
                                               pairwise_passes=passes))
    one = hierarchy(tmp_path, rp, ci, v, passes, 1)
    many = hierarchy(tmp_path, rp, ci, v, passes, 7)
    nl, aggt, mats = one
    assert nl == Mo.nlevels() >= 3
    for l in range(nl - 1):
        np.testing.assert_array_equal(aggt[l], Mo.aggregates(l))
        np.testing.assert_array_equal(many[1][l], aggt[l])
    for l in range(1, nl):
        orp, oci, ov = Mo.level(l)
        for got, other, want in zip(mats[l], many[2][l], (orp, oci, ov)):
            np.testing.assert_array_equal(got, want)
            np.testing.assert_array_equal(other, want)
