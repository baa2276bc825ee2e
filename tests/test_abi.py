"""The C ABI library loads and exports every function include/fsx.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

from flecsolve_amd import fsx

ROOT = Path(__file__).resolve().parents[1]


def declared(header):
    txt = (ROOT / "include" / header).read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fsx_[a-z0-9_]+)\s*\(", txt)))


def test_libfsx_exports_every_declared_symbol():
    L = ctypes.CDLL(str(fsx.LIB_PATH))
    names = declared("fsx.h")
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    names = set(declared("fsx.h"))
    assert names <= set(fsx.exported_symbols()) | {"fsx_last_error"}


def test_problems_library_exports():
    from flecsolve_amd import problems
    L = ctypes.CDLL(str(problems.LIB_PATH))
    missing = [n for n in declared("fsx_problems.h") if not hasattr(L, n)]
    assert not missing, missing


def test_error_path_without_gpu():
    # creating a context on a machine without the device must fail cleanly
    import pytest
    from tests.conftest import _gpu_present
    if _gpu_present():
        pytest.skip("GPU present")
    h = ctypes.c_void_p()
    rc = fsx.lib().fsx_ctx_create(0, ctypes.byref(h))
    assert rc != 0
    assert fsx.lib().fsx_last_error()
