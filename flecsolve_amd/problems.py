"""Deterministic synthetic problems (binding of libfsx_problems.so).

All generators return numpy arrays (int32 CSR indices, float64 values) and
are pure functions of their arguments and seeds.  Grid conventions are in
include/fsx_problems.h.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libfsx_problems.so"
SEEDS_PATH = _PKG / "seeds.json"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not built; run `python -m flecsolve_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        L.fsx_gen_poisson2d_nnz.restype = C.c_int64
        L.fsx_gen_poisson2d_nnz.argtypes = [C.c_int]
        L.fsx_gen_poisson2d.argtypes = [C.c_int, P, P, P]
        L.fsx_gen_grf_lognormal.argtypes = [C.c_int, C.c_double, C.c_double,
                                            C.c_uint64, P]
        L.fsx_gen_diffusion3d_nnz.restype = C.c_int64
        L.fsx_gen_diffusion3d_nnz.argtypes = [C.c_int]
        L.fsx_gen_diffusion3d.argtypes = [C.c_int, P, C.c_double, C.c_double, P, P, P]
        L.fsx_gen_convdiff3d.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                         P, P, P]
        L.fsx_gen_uniform.argtypes = [P, C.c_int64, C.c_uint64, C.c_double, C.c_double]
        L.fsx_gen_bumps.argtypes = [C.c_int, C.c_int, C.c_uint64, P]
        L.fsx_gen_batch_sizes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, P, P, P]
        L.fsx_gen_batch.argtypes = [C.c_int, P, P, P, C.c_double, C.c_double,
                                    C.c_uint64, P, P, P]
        _lib = L
    return _lib


def seeds() -> dict:
    return json.loads(SEEDS_PATH.read_text())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _alloc(nrows, nnz):
    return (np.empty(nrows + 1, np.int32), np.empty(nnz, np.int32),
            np.empty(nnz, np.float64))


def poisson2d(n: int):
    """5-point [4,-1] Laplacian on an n x n interior grid."""
    rp, ci, v = _alloc(n * n, lib().fsx_gen_poisson2d_nnz(n))
    assert lib().fsx_gen_poisson2d(n, _p(rp), _p(ci), _p(v)) == 0
    return rp, ci, v


def lognormal_field(n: int, sigma: float, ell: float, seed: int) -> np.ndarray:
    """k = exp(sigma*g) on the (n+2)^3 lattice (g: unit Gaussian random field)."""
    k = np.empty((n + 2) ** 3, np.float64)
    assert lib().fsx_gen_grf_lognormal(n, sigma, ell, seed, _p(k)) == 0
    return k


def diffusion3d(n: int, k: np.ndarray, scale: float = 1.0, shift: float = 0.0):
    """shift*I + scale*L_k, 7-point, arithmetic-mean face coefficients."""
    k = np.ascontiguousarray(k, np.float64)
    assert k.size == (n + 2) ** 3
    rp, ci, v = _alloc(n ** 3, lib().fsx_gen_diffusion3d_nnz(n))
    assert lib().fsx_gen_diffusion3d(n, _p(k), scale, shift, _p(rp), _p(ci), _p(v)) == 0
    return rp, ci, v


def convdiff3d(n: int, c=(10.0, 5.0, 2.5)):
    """Upwind convection-diffusion with cell velocities c (x, y, z)."""
    rp, ci, v = _alloc(n ** 3, lib().fsx_gen_diffusion3d_nnz(n))
    assert lib().fsx_gen_convdiff3d(n, c[0], c[1], c[2], _p(rp), _p(ci), _p(v)) == 0
    return rp, ci, v


def uniform(n: int, seed: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    x = np.empty(n, np.float64)
    assert lib().fsx_gen_uniform(_p(x), n, seed, lo, hi) == 0
    return x


def bumps(n: int, nb: int, seed: int) -> np.ndarray:
    u = np.empty(n ** 3, np.float64)
    assert lib().fsx_gen_bumps(n, nb, seed, _p(u)) == 0
    return u


def batch(nsys: int, emin: int, emax: int, seed_e: int, seed_k: int,
          sigma: float = 1.0, ell: float = 4.0):
    """Block-diagonal batch of 3D lognormal-k diffusion systems.
    Returns (edges, sys_rows, rowptr, colind, vals)."""
    edges = np.empty(nsys, np.int32)
    rows = np.empty(nsys + 1, np.int64)
    nnzs = np.empty(nsys + 1, np.int64)
    assert lib().fsx_gen_batch_sizes(nsys, emin, emax, seed_e, _p(edges), _p(rows),
                                     _p(nnzs)) == 0
    rp, ci, v = _alloc(int(rows[-1]), int(nnzs[-1]))
    assert lib().fsx_gen_batch(nsys, _p(edges), _p(rows), _p(nnzs), sigma, ell, seed_k,
                               _p(rp), _p(ci), _p(v)) == 0
    return edges, rows, rp, ci, v
