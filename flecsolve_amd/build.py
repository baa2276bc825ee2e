"""Build the native libraries in-tree.

    python -m flecsolve_amd.build            # everything
    python -m flecsolve_amd.build --no-oracle

Outputs (git-ignored, shipped with the tree):
    flecsolve_amd/libfsx.so           HIP kernels + C ABI (hipcc, gfx950)
    flecsolve_amd/libfsx_problems.so  synthetic problem generators (g++)
    oracle/liboracle.so               CPU checker (built by oracle/Makefile)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INC = ROOT / "include"
BUILD = PKG / "_build"

ARCH = os.environ.get("FSX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["fsx_core.hip", "fsx_spmv.hip", "fsx_cg.hip", "fsx_batched.hip",
               "fsx_adjoint.hip", "fsx_smooth.hip", "fsx_amg.hip", "fsx_parcsr.hip"]
CPP_SOURCES = ["fsx_krylov.cpp", "fsx_io.cpp"]

# -ffp-contract=off: no fused multiply-add contraction, so elementwise results
# round exactly like the reference's serial loops.  Packed-FP32 instructions
# are disabled for all device code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", str(INC), "-I", str(CSRC),
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + r.stdout)
    # hipcc's host pass warns about the device-only target feature: harmless
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out, file=sys.stderr)


def _compile(src: Path, obj: Path) -> Path:
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        # host-only C++ above the C ABI: plain g++
        cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
               "-Wall", "-I", str(INC), "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_fsx(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in HIP_SOURCES + CPP_SOURCES]
    objs = [BUILD / (s.stem + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, srcs, objs))
    out = PKG / "libfsx.so"
    tmp = PKG / "libfsx.so.tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp),
          *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_problems() -> Path:
    out = PKG / "libfsx_problems.so"
    tmp = PKG / "libfsx_problems.so.tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
          "-I", str(INC), str(CSRC / "problems.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_config_check() -> Path:
    """Config-file reading check (tests/cpp/config_check.cpp; no device use)."""
    src = ROOT / "tests" / "cpp" / "config_check.cpp"
    out = ROOT / "tests" / "cpp" / "config_check"
    tmp = ROOT / "tests" / "cpp" / "config_check.tmp"
    _run(["g++", "-O1", "-std=c++17", "-Wall", "-I", str(INC), str(src), "-o", str(tmp),
          "-L", str(PKG), "-lfsx", f"-Wl,-rpath,{PKG}",
          "-Wl,-rpath,$ORIGIN/../../flecsolve_amd"])
    os.replace(tmp, out)
    return out


def build_cpp_tests() -> Path:
    """C++ drop-in check against the header API (tests/cpp)."""
    src = ROOT / "tests" / "cpp" / "cg_dropin.cpp"
    out = ROOT / "tests" / "cpp" / "cg_dropin"
    tmp = ROOT / "tests" / "cpp" / "cg_dropin.tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", str(INC), str(src), "-o", str(tmp),
          "-L", str(PKG), "-lfsx", "-lfsx_problems", f"-Wl,-rpath,{PKG}",
          "-Wl,-rpath,$ORIGIN/../../flecsolve_amd"])
    os.replace(tmp, out)
    return out


def build_amg_setup_check() -> Path:
    """UA-AMG setup on the host alone (tests/cpp/amg_setup_check.cpp over
    flecsolve_amd/csrc/fsx_amg_setup.hpp, the code fsx_amg_create runs)."""
    src = ROOT / "tests" / "cpp" / "amg_setup_check.cpp"
    out = ROOT / "tests" / "cpp" / "amg_setup_check"
    tmp = ROOT / "tests" / "cpp" / "amg_setup_check.tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-ffp-contract=off", "-pthread",
          "-I", str(CSRC), str(src), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_bdf_checks() -> None:
    """Time-integrator check (tests/cpp/bdf_check.cpp) over host vectors (no
    device use) and over device vectors, and the heat example
    (flecsolve_amd/examples/heat_implicit.cpp)."""
    cpp = ROOT / "tests" / "cpp"
    common = ["g++", "-O2", "-std=c++17", "-Wall", "-ffp-contract=off", "-I", str(INC),
              "-I", str(cpp)]
    rpath = [f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../flecsolve_amd"]
    for out, extra, libs in (("bdf_check_host", ["-DBDF_HOST"], ["-lfsx_problems"]),
                             ("bdf_check", [], ["-lfsx", "-lfsx_problems"])):
        tmp = cpp / (out + ".tmp")
        _run(common + extra + [str(cpp / "bdf_check.cpp"), "-o", str(tmp), "-L", str(PKG)]
             + libs + rpath)
        os.replace(tmp, cpp / out)
    tmp = PKG / "heat_implicit.tmp"
    _run(common + [str(PKG / "examples" / "heat_implicit.cpp"), "-o", str(tmp), "-L", str(PKG),
                   "-lfsx", "-lfsx_problems", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, PKG / "heat_implicit")


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


def build_all(oracle: bool = True) -> None:
    build_problems()
    if oracle:
        build_oracle()
    build_fsx()
    build_cpp_tests()
    build_config_check()
    build_amg_setup_check()
    build_bdf_checks()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    if a.clean and BUILD.exists():
        shutil.rmtree(BUILD)
    build_all(oracle=not a.no_oracle)
    print("built:", PKG / "libfsx.so")


if __name__ == "__main__":
    main()
