"""flecsolve_amd -- MI355X-native Krylov solve path of lanl/flecsolve.

Native pieces (built in-tree by ``python -m flecsolve_amd.build``):
  libfsx.so           HIP (gfx950) kernels and the C ABI of include/fsx.h,
                      plus the generic C++ solvers of include/flecsolve/
  libfsx_problems.so  deterministic synthetic problem generators

The Python modules are thin ctypes bindings used by tests and benchmarks.
"""
from . import fsx  # noqa: F401

__all__ = ["fsx"]
