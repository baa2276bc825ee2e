"""ctypes binding of libfsx (include/fsx.h) for tests and benchmarks.

The library must have been built in-tree (``python -m flecsolve_amd.build``);
importing this module without it raises -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
# FSX_LIB overrides the library path (e.g. to load another build of libfsx)
LIB_PATH = Path(os.environ["FSX_LIB"]) if os.environ.get("FSX_LIB") else _PKG / "libfsx.so"

F64, F32 = 0, 1
I32, I64 = 0, 1

STOP_REASONS = ["converged_atol", "converged_rtol", "converged_user",
                "diverged_dtol", "diverged_iters", "diverged_breakdown",
                "unknown"]


class Settings(C.Structure):
    _fields_ = [("maxiter", C.c_int), ("rtol", C.c_float), ("atol", C.c_float),
                ("use_zero_guess", C.c_int), ("max_krylov_dim", C.c_int),
                ("pre_side", C.c_int), ("restart", C.c_int)]

    def __init__(self, maxiter=100, rtol=0.0, atol=0.0, use_zero_guess=False,
                 max_krylov_dim=-1, pre_side=1, restart=False):
        super().__init__(maxiter, rtol, atol, int(use_zero_guess),
                         max_krylov_dim, pre_side, int(restart))


class Info(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("restarts", C.c_int),
                ("res_norm_initial", C.c_float), ("res_norm_final", C.c_float),
                ("sol_norm_initial", C.c_float), ("sol_norm_final", C.c_float),
                ("rhs_norm", C.c_float)]

    @property
    def reason(self) -> str:
        return STOP_REASONS[self.status]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class IRInfo(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("inner_iters_total", C.c_int),
                ("status", C.c_int), ("rel_res_final", C.c_double),
                ("rhs_norm", C.c_double)]


class BEStats(C.Structure):
    _fields_ = [("forward_solves", C.c_int), ("recompute_solves", C.c_int),
                ("adjoint_solves", C.c_int), ("failed_step", C.c_int),
                ("forward_iters", C.c_int64), ("recompute_iters", C.c_int64),
                ("adjoint_iters", C.c_int64), ("checkpoint_bytes", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None

_P = C.c_void_p
# fsx_diag_fn: int (*)(void * user, fsx_vec * x, double res_norm)
DIAG_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_double)
_SIGS = {
    "fsx_last_error": (C.c_char_p, []),
    "fsx_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "fsx_ctx_destroy": (C.c_int, [_P]),
    "fsx_sync": (C.c_int, [_P]),
    "fsx_ctx_stream": (C.c_int, [_P, C.POINTER(_P)]),
    "fsx_ctx_counters": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "fsx_alloc_stats": (C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "fsx_time_cg_spmv": (C.c_int, [_P, _P, _P, _P, C.c_int, C.POINTER(C.c_double)]),
    "fsx_time_cg_spmv_sr": (C.c_int, [_P, _P, _P, _P, C.c_int, C.POINTER(C.c_double)]),
    "fsx_csr_layout_bytes": (C.c_int, [_P, C.c_size_t, C.POINTER(C.c_uint64)]),
    "fsx_vec_create": (C.c_int, [_P, C.c_size_t, C.c_int, C.POINTER(_P)]),
    "fsx_vec_wrap": (C.c_int, [_P, _P, C.c_size_t, C.c_int, C.POINTER(_P)]),
    "fsx_vec_retain": (C.c_int, [_P]),
    "fsx_vec_release": (C.c_int, [_P]),
    "fsx_vec_size": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "fsx_vec_dtype": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "fsx_vec_id": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "fsx_vec_device_ptr": (C.c_int, [_P, C.POINTER(_P)]),
    "fsx_vec_upload": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "fsx_vec_download": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "fsx_vec_get": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_double)]),
    "fsx_vec_set": (C.c_int, [_P, _P, C.c_size_t, C.c_double]),
    "fsx_copy": (C.c_int, [_P, _P, _P]),
    "fsx_zero": (C.c_int, [_P, _P]),
    "fsx_set_scalar": (C.c_int, [_P, C.c_double, _P]),
    "fsx_scale": (C.c_int, [_P, C.c_double, _P]),
    "fsx_scale_to": (C.c_int, [_P, C.c_double, _P, _P]),
    "fsx_add": (C.c_int, [_P, _P, _P, _P]),
    "fsx_subtract": (C.c_int, [_P, _P, _P, _P]),
    "fsx_multiply": (C.c_int, [_P, _P, _P, _P]),
    "fsx_divide": (C.c_int, [_P, _P, _P, _P]),
    "fsx_reciprocal": (C.c_int, [_P, _P, _P]),
    "fsx_linear_sum": (C.c_int, [_P, C.c_double, _P, C.c_double, _P, _P]),
    "fsx_axpy": (C.c_int, [_P, C.c_double, _P, _P, _P]),
    "fsx_axpby": (C.c_int, [_P, C.c_double, C.c_double, _P, _P]),
    "fsx_abs": (C.c_int, [_P, _P, _P]),
    "fsx_add_scalar": (C.c_int, [_P, _P, C.c_double, _P]),
    "fsx_set_random": (C.c_int, [_P, _P, C.c_uint]),
    "fsx_convert": (C.c_int, [_P, _P, _P, C.c_double]),
    "fsx_dot": (C.c_int, [_P, _P, _P, C.POINTER(C.c_double)]),
    "fsx_nrm1": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "fsx_nrm2": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "fsx_nrminf": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "fsx_nrmp": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_double)]),
    "fsx_min": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "fsx_max": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "fsx_csr_create": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_int64, C.c_int,
                                 _P, _P, C.c_int, _P, C.POINTER(_P)]),
    "fsx_csr_convert": (C.c_int, [_P, _P, C.c_int, C.POINTER(_P)]),
    "fsx_csr_destroy": (C.c_int, [_P]),
    "fsx_csr_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                               C.POINTER(C.c_int64), C.POINTER(C.c_int),
                               C.POINTER(C.c_int)]),
    "fsx_csr_plan": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fsx_csr_set_plan": (C.c_int, [_P, C.c_int, C.c_int]),
    "fsx_csr_download": (C.c_int, [_P, _P, _P, _P, _P]),
    "fsx_csr_spmv": (C.c_int, [_P, _P, _P, _P]),
    "fsx_csr_residual": (C.c_int, [_P, _P, _P, _P, _P]),
    "fsx_csr_diag_inv": (C.c_int, [_P, _P, _P]),
    "fsx_csr_spmv_t": (C.c_int, [_P, _P, _P, _P]),
    "fsx_parcsr_create": (C.c_int, [_P, _P, C.c_int, C.POINTER(_P)]),
    "fsx_parcsr_destroy": (C.c_int, [_P]),
    "fsx_parcsr_colour_info": (C.c_int, [_P, C.c_int, _P]),
    "fsx_parcsr_download": (C.c_int, [_P, _P, C.c_int, _P, _P, _P, _P, _P, _P, _P]),
    "fsx_parcsr_spmv": (C.c_int, [_P, _P, _P, _P]),
    "fsx_parcsr_residual": (C.c_int, [_P, _P, _P, _P, _P]),
    "fsx_parcsr_diag": (C.c_int, [_P, C.POINTER(_P)]),
    "fsx_time_parcsr_spmv": (C.c_int, [_P, _P, _P, _P, C.c_int, C.POINTER(C.c_double)]),
    "fsx_spmv_dot": (C.c_int, [_P, _P, _P, _P, C.POINTER(C.c_double)]),
    "fsx_jacobi_apply": (C.c_int, [_P, _P, _P, _P]),
    "fsx_jacobi_relax": (C.c_int, [_P, _P, C.c_float, C.c_int, _P, _P]),
    "fsx_gs_create": (C.c_int, [_P, _P, C.c_int, C.POINTER(_P)]),
    "fsx_gs_relax": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P]),
    "fsx_gs_levels": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fsx_gs_destroy": (C.c_int, [_P]),
    "fsx_amg_create": (C.c_int, [_P, _P, _P, C.POINTER(_P)]),
    "fsx_amg_levels": (C.c_int, [_P, C.POINTER(C.c_int), _P, _P, C.c_int]),
    "fsx_amg_aggregates": (C.c_int, [_P, C.c_int, _P]),
    "fsx_amg_level_matrix": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "fsx_amg_apply": (C.c_int, [_P, _P, _P, _P, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "fsx_amg_destroy": (C.c_int, [_P]),
    "fsx_krylov_cg_amg": (C.c_int, [_P, _P, _P, C.c_int, _P, _P, _P, _P, _P, C.c_int]),
    "fsx_cg_solve": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(Settings),
                               C.POINTER(Info)]),
    "fsx_bicgstab_solve": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(Settings),
                                     C.POINTER(Info)]),
    "fsx_gmres_solve": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(Settings),
                                  C.POINTER(Info)]),
    "fsx_cg_solve_noalloc": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(Settings),
                                       C.POINTER(Info)]),
    "fsx_cg_solve_diag": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(Settings), DIAG_FN, _P,
                                    C.POINTER(Info)]),
    "fsx_cg_history": (C.c_int, [_P, C.POINTER(C.c_double), C.c_int,
                                 C.POINTER(C.c_int)]),
    "fsx_ir_cg_solve": (C.c_int, [_P, _P, _P, C.c_int, _P, _P, C.c_double,
                                  C.c_int, C.POINTER(Settings),
                                  C.POINTER(IRInfo)]),
    "fsx_ir_cg_rr_solve": (C.c_int, [_P, _P, _P, _P, _P, C.c_double, C.c_int,
                                     C.POINTER(Settings), C.POINTER(IRInfo)]),
    "fsx_krylov_cg": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P,
                                C.POINTER(Settings), C.POINTER(Info),
                                C.POINTER(C.c_double), C.c_int]),
    "fsx_krylov_cg_generic": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P,
                                        C.POINTER(Settings), C.POINTER(Info),
                                        C.POINTER(C.c_double), C.c_int]),
    "fsx_krylov_fcg": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P,
                                 C.POINTER(Settings), C.POINTER(Info)]),
    "fsx_krylov_gmres": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P,
                                   C.POINTER(Settings), C.POINTER(Info),
                                   C.POINTER(C.c_double), C.c_int]),
    "fsx_krylov_bicgstab": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P,
                                      C.POINTER(Settings), C.POINTER(Info),
                                      C.POINTER(C.c_double), C.c_int]),
    "fsx_batch_create": (C.c_int, [_P, _P, C.c_int64, _P, C.POINTER(_P)]),
    "fsx_batch_destroy": (C.c_int, [_P]),
    "fsx_batch_config": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]),
    "fsx_batch_set_config": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "fsx_batch_tune": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, C.c_int,
                                 C.POINTER(C.c_double)]),
    "fsx_batched_cg": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(Settings), _P, _P, _P]),
    "fsx_batched_ir_cg": (C.c_int, [_P, _P, _P, C.c_int, _P, _P, C.c_double, C.c_int,
                                    C.POINTER(Settings), _P, _P, _P, _P]),
    "fsx_batched_rr_cg": (C.c_int, [_P, _P, _P, _P, _P, C.c_double, C.c_int,
                                    C.POINTER(Settings), _P, _P, _P, _P]),
    "fsx_mm_read_header": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int64)]),
    "fsx_mm_read_coo": (C.c_int, [C.c_char_p, C.c_int64, _P, _P, _P, C.POINTER(C.c_int64)]),
    "fsx_coo_tocsr": (C.c_int, [C.c_int64, C.c_int64, _P, _P, _P, _P, _P, _P]),
    "fsx_csr_from_mm": (C.c_int, [_P, C.c_char_p, C.c_int, C.POINTER(_P)]),
    "fsx_csr_transpose": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "fsx_csr_set_values": (C.c_int, [_P, _P, _P]),
    "fsx_csr_set_values_f32_from_f64": (C.c_int, [_P, _P, _P]),
    "fsx_csr_shift_scale": (C.c_int, [_P, _P, C.c_double, C.c_double, _P]),
    "fsx_csr_set_values_f32_sym_scaled": (C.c_int, [_P, _P, _P]),
    "fsx_reduce_async": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "fsx_cg_update": (C.c_int, [_P, C.c_double, _P, _P, _P, _P, _P, _P,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "fsx_csr_sampled_outer": (C.c_int, [_P, _P, _P, _P, C.c_double, _P]),
    "fsx_diffusion3d_values": (C.c_int, [_P, C.c_int, _P, C.c_double, C.c_double, _P]),
    "fsx_diffusion3d_coeff_grad": (C.c_int, [_P, C.c_int, _P, _P, _P, C.c_double, _P]),
    "fsx_be_forward": (C.c_int, [_P, _P, _P, _P, C.c_int, C.POINTER(Settings), _P,
                                 C.POINTER(BEStats)]),
    "fsx_be_adjoint": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_int, C.c_int,
                                 C.POINTER(Settings), C.POINTER(C.c_double), _P, _P, _P,
                                 C.POINTER(BEStats)]),
    "fsx_be_forward_ir": (C.c_int, [_P, _P, _P, _P, C.c_int, C.POINTER(Settings),
                                    C.POINTER(Settings), _P, C.POINTER(BEStats)]),
    "fsx_be_adjoint_ir": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, C.c_int,
                                    C.POINTER(Settings), C.POINTER(Settings),
                                    C.POINTER(C.c_double), _P, _P, _P, C.POINTER(BEStats)]),
    "fsx_be_forward_rr": (C.c_int, [_P, _P, _P, _P, C.c_int, C.POINTER(Settings),
                                    C.POINTER(Settings), _P, C.POINTER(BEStats)]),
    "fsx_be_adjoint_rr": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, C.c_int,
                                    C.POINTER(Settings), C.POINTER(Settings),
                                    C.POINTER(C.c_double), _P, _P, _P, C.POINTER(BEStats)]),
    "fsx_time_spmv": (C.c_int, [_P, _P, _P, _P, C.c_int, C.POINTER(C.c_double)]),
    "fsx_time_copy": (C.c_int, [_P, _P, _P, C.c_int, C.POINTER(C.c_double)]),
}


class FsxError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libfsx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FsxError(f"{LIB_PATH} not built; run `python -m flecsolve_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().fsx_last_error().decode()
        raise FsxError(f"{what} failed (status {rc}): {msg}")


def _ptr(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


class Context:
    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        check(lib().fsx_ctx_create(device, C.byref(self.h)), "fsx_ctx_create")

    def sync(self):
        check(lib().fsx_sync(self.h), "fsx_sync")

    def stream(self) -> int:
        s = C.c_void_p()
        check(lib().fsx_ctx_stream(self.h, C.byref(s)), "fsx_ctx_stream")
        return s.value or 0

    def counters(self) -> tuple[int, int]:
        a, b = C.c_uint64(), C.c_uint64()
        check(lib().fsx_ctx_counters(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            check(lib().fsx_ctx_destroy(self.h), "fsx_ctx_destroy")
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # convenience constructors
    def vec(self, data, dtype=None) -> "Vector":
        a = np.asarray(data)
        dt = F32 if (dtype == F32 or (dtype is None and a.dtype == np.float32)) else F64
        v = Vector(self, a.size, dt)
        v.upload(a)
        return v

    def zeros(self, n: int, dtype: int = F64) -> "Vector":
        return Vector(self, n, dtype)

    def csr(self, rowptr, colind, vals, ncols=None, vtype=None) -> "CSR":
        return CSR(self, rowptr, colind, vals, ncols=ncols, vtype=vtype)


class Vector:
    def __init__(self, ctx: Context, n: int, dtype: int = F64, _handle=None):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = dtype
        if _handle is not None:
            self.h = _handle
        else:
            self.h = C.c_void_p()
            check(lib().fsx_vec_create(ctx.h, self.n, dtype, C.byref(self.h)),
                  "fsx_vec_create")

    @property
    def np_dtype(self):
        return np.float64 if self.dtype == F64 else np.float32

    def upload(self, a) -> "Vector":
        a = np.ascontiguousarray(a, dtype=self.np_dtype).ravel()
        check(lib().fsx_vec_upload(self.ctx.h, self.h, _ptr(a), a.size), "upload")
        return self

    def __getitem__(self, i: int) -> float:
        """One element (fsx_vec_get: vec::core::operator[] read)."""
        out = C.c_double()
        check(lib().fsx_vec_get(self.ctx.h, self.h, int(i), C.byref(out)), "fsx_vec_get")
        return out.value

    def __setitem__(self, i: int, value: float) -> None:
        """One element (fsx_vec_set: vec::core::operator[] write)."""
        check(lib().fsx_vec_set(self.ctx.h, self.h, int(i), float(value)), "fsx_vec_set")

    def get(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.np_dtype)
        check(lib().fsx_vec_download(self.ctx.h, self.h, _ptr(out), self.n),
              "download")
        return out

    def id(self) -> int:
        i = C.c_uint64()
        check(lib().fsx_vec_id(self.h, C.byref(i)))
        return i.value

    def release(self):
        if self.h:
            check(lib().fsx_vec_release(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class CSR:
    def __init__(self, ctx: Context, rowptr, colind, vals, ncols=None, vtype=None,
                 _handle=None):
        self.ctx = ctx
        if _handle is not None:
            self.h = _handle
            nr, nc, nz = C.c_int64(), C.c_int64(), C.c_int64()
            it, vt = C.c_int(), C.c_int()
            check(lib().fsx_csr_info(self.h, C.byref(nr), C.byref(nc), C.byref(nz),
                                     C.byref(it), C.byref(vt)))
            self.nrows, self.ncols, self.nnz = nr.value, nc.value, nz.value
            self.itype, self.vtype = it.value, vt.value
            return
        rowptr = np.ascontiguousarray(rowptr)
        it = I32 if rowptr.dtype == np.int32 else I64
        idt = np.int32 if it == I32 else np.int64
        rowptr = rowptr.astype(idt, copy=False)
        colind = np.ascontiguousarray(colind, dtype=idt)
        if vtype is None:
            vtype = F32 if np.asarray(vals).dtype == np.float32 else F64
        vals = np.ascontiguousarray(vals, dtype=np.float64 if vtype == F64 else np.float32)
        self.nrows = rowptr.size - 1
        self.ncols = self.nrows if ncols is None else int(ncols)
        self.nnz = int(colind.size)
        self.itype, self.vtype = it, vtype
        self.h = C.c_void_p()
        check(lib().fsx_csr_create(ctx.h, self.nrows, self.ncols, self.nnz, it,
                                   _ptr(rowptr), _ptr(colind), vtype, _ptr(vals),
                                   C.byref(self.h)), "fsx_csr_create")

    def convert(self, vtype: int) -> "CSR":
        h = C.c_void_p()
        check(lib().fsx_csr_convert(self.ctx.h, self.h, vtype, C.byref(h)),
              "fsx_csr_convert")
        return CSR(self.ctx, None, None, None, _handle=h)

    def layout_bytes(self, x_bytes: int) -> int:
        """Compulsory bytes of one product on the layout this matrix's plan reads."""
        out = C.c_uint64()
        check(lib().fsx_csr_layout_bytes(self.h, x_bytes, C.byref(out)), "layout_bytes")
        return out.value

    def plan(self) -> tuple[int, int]:
        k, l = C.c_int(), C.c_int()
        check(lib().fsx_csr_plan(self.h, C.byref(k), C.byref(l)))
        return k.value, l.value

    def set_plan(self, kind: int, lanes: int = 0):
        check(lib().fsx_csr_set_plan(self.h, kind, lanes), "fsx_csr_set_plan")

    def download(self):
        idt = np.int32 if self.itype == I32 else np.int64
        vdt = np.float64 if self.vtype == F64 else np.float32
        rp = np.empty(self.nrows + 1, idt)
        ci = np.empty(self.nnz, idt)
        vv = np.empty(self.nnz, vdt)
        check(lib().fsx_csr_download(self.ctx.h, self.h, _ptr(rp), _ptr(ci), _ptr(vv)))
        return rp, ci, vv

    def spmv(self, x: Vector, y: Vector):
        check(lib().fsx_csr_spmv(self.ctx.h, self.h, x.h, y.h), "fsx_csr_spmv")

    def spmv_t(self, x: Vector, y: Vector):
        check(lib().fsx_csr_spmv_t(self.ctx.h, self.h, x.h, y.h), "fsx_csr_spmv_t")

    def residual(self, b: Vector, x: Vector, r: Vector):
        check(lib().fsx_csr_residual(self.ctx.h, self.h, b.h, x.h, r.h), "residual")

    def set_values(self, vals):
        v = np.ascontiguousarray(vals, np.float64 if self.vtype == F64 else np.float32)
        assert v.size == self.nnz
        check(lib().fsx_csr_set_values(self.ctx.h, self.h, _ptr(v)), "fsx_csr_set_values")

    def set_values_from(self, A64: "CSR"):
        """FP32 copy takes A64's current values (same structure)."""
        check(lib().fsx_csr_set_values_f32_from_f64(self.ctx.h, self.h, A64.h),
              "fsx_csr_set_values_f32_from_f64")

    def set_values_sym_scaled(self, A64: "CSR"):
        """FP32 copy takes S A64 S, S = D^{-1/2} (inner matrix of
        ir_cg_solve(..., jacobi=2))."""
        check(lib().fsx_csr_set_values_f32_sym_scaled(self.ctx.h, self.h, A64.h),
              "fsx_csr_set_values_f32_sym_scaled")

    def shift_scale(self, F: "CSR", alpha: float, beta: float):
        """self = alpha I + beta F on F's pattern (self = F.convert(F64))."""
        check(lib().fsx_csr_shift_scale(self.ctx.h, F.h, alpha, beta, self.h),
              "fsx_csr_shift_scale")

    def transpose(self) -> "CSR":
        h = C.c_void_p()
        check(lib().fsx_csr_transpose(self.ctx.h, self.h, C.byref(h)), "fsx_csr_transpose")
        return CSR(self.ctx, None, None, None, _handle=h)

    def sampled_outer(self, lam: Vector, x: Vector, scale: float, out: Vector):
        """out[e] += (scale * lam[row e]) * x[col e] on the stored entries."""
        check(lib().fsx_csr_sampled_outer(self.ctx.h, self.h, lam.h, x.h, scale, out.h),
              "fsx_csr_sampled_outer")

    def set_diffusion3d_values(self, n: int, k: Vector, shift: float, scale: float):
        """Values <- shift I + scale L_k (7-point, arithmetic-mean faces)."""
        check(lib().fsx_diffusion3d_values(self.ctx.h, n, k.h, shift, scale, self.h),
              "fsx_diffusion3d_values")

    def diffusion3d_coeff_grad(self, n: int, G: Vector, k: Vector, scale: float) -> Vector:
        g = Vector(self.ctx, (n + 2) ** 3)
        check(lib().fsx_diffusion3d_coeff_grad(self.ctx.h, n, self.h, G.h, k.h, scale, g.h),
              "fsx_diffusion3d_coeff_grad")
        return g

    def diag_inv(self, dtype: int = F64) -> Vector:
        d = Vector(self.ctx, self.nrows, dtype)
        check(lib().fsx_csr_diag_inv(self.ctx.h, self.h, d.h), "diag_inv")
        return d

    def destroy(self):
        if self.h and getattr(self, "_owner", None) is None:
            check(lib().fsx_csr_destroy(self.h))
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---- free functions --------------------------------------------------------

def _red(name, ctx, *vs) -> float:
    out = C.c_double()
    check(getattr(lib(), name)(ctx.h, *[v.h for v in vs], C.byref(out)), name)
    return out.value


def dot(x: Vector, y: Vector) -> float:
    return _red("fsx_dot", x.ctx, x, y)


def nrm2(x: Vector) -> float:
    return _red("fsx_nrm2", x.ctx, x)


def nrm1(x: Vector) -> float:
    return _red("fsx_nrm1", x.ctx, x)


def nrminf(x: Vector) -> float:
    return _red("fsx_nrminf", x.ctx, x)


def nrmp(x: Vector, p: int) -> float:
    """lp_norm<p> of the FleCSI vector path: (sum x_i^p)^(1/p), no abs."""
    out = C.c_double()
    check(lib().fsx_nrmp(x.ctx.h, x.h, int(p), C.byref(out)), "fsx_nrmp")
    return out.value


def vmin(x: Vector) -> float:
    return _red("fsx_min", x.ctx, x)


def vmax(x: Vector) -> float:
    return _red("fsx_max", x.ctx, x)


def spmv_dot(A: CSR, p: Vector, w: Vector) -> float:
    out = C.c_double()
    check(lib().fsx_spmv_dot(A.ctx.h, A.h, p.h, w.h, C.byref(out)), "spmv_dot")
    return out.value


def cg_solve(A: CSR, b: Vector, x: Vector, settings: Settings,
             dinv: Vector | None = None) -> Info:
    info = Info()
    check(lib().fsx_cg_solve(A.ctx.h, A.h, dinv.h if dinv is not None else None,
                             b.h, x.h, C.byref(settings), C.byref(info)),
          "fsx_cg_solve")
    return info


def cg_solve_diag(A: CSR, b: Vector, x: Vector, settings: Settings, diag,
                  dinv: Vector | None = None) -> Info:
    """fsx_cg_solve_diag: the fused CG calling diag(x, res) on the host after
    every iteration (before the tolerance test); a true return stops the
    solve with converged_user.  An exception in diag aborts the solve and is
    raised here."""
    err = []

    def _cb(_user, _xh, res):
        try:
            return 1 if diag(x, res) else 0
        except Exception as e:  # carried back to the caller
            err.append(e)
            return -1

    fn = DIAG_FN(_cb)
    info = Info()
    rc = lib().fsx_cg_solve_diag(A.ctx.h, A.h, dinv.h if dinv is not None else None,
                                 b.h, x.h, C.byref(settings), fn, None, C.byref(info))
    if err:
        raise err[0]
    check(rc, "fsx_cg_solve_diag")
    return info


def cg_solve_noalloc(A: CSR, b: Vector, x: Vector, settings: Settings,
                     dinv: Vector | None = None) -> Info:
    """fsx_cg_solve_noalloc: the fused CG with every work buffer in the
    library's static per-device scratch array (n <= NOALLOC_MAX_ROWS)."""
    info = Info()
    check(lib().fsx_cg_solve_noalloc(A.ctx.h, A.h, dinv.h if dinv is not None else None,
                                     b.h, x.h, C.byref(settings), C.byref(info)),
          "fsx_cg_solve_noalloc")
    return info


NOALLOC_MAX_ROWS = 1 << 20


def alloc_stats() -> tuple[int, int]:
    """(hipMalloc calls the library has made, bytes its live allocations hold)."""
    n, b = C.c_uint64(), C.c_uint64()
    check(lib().fsx_alloc_stats(C.byref(n), C.byref(b)), "fsx_alloc_stats")
    return n.value, b.value


def bicgstab_solve(A: CSR, b: Vector, x: Vector, settings: Settings,
                   dinv: Vector | None = None) -> Info:
    info = Info()
    check(lib().fsx_bicgstab_solve(A.ctx.h, A.h, dinv.h if dinv is not None else None,
                                   b.h, x.h, C.byref(settings), C.byref(info)),
          "fsx_bicgstab_solve")
    return info


def gmres_solve(A: CSR, b: Vector, x: Vector, settings: Settings,
                dinv: Vector | None = None) -> Info:
    info = Info()
    check(lib().fsx_gmres_solve(A.ctx.h, A.h, dinv.h if dinv is not None else None,
                                b.h, x.h, C.byref(settings), C.byref(info)),
          "fsx_gmres_solve")
    return info


def cg_history(ctx: Context) -> np.ndarray:
    cnt = C.c_int()
    check(lib().fsx_cg_history(ctx.h, None, 0, C.byref(cnt)))
    out = np.empty(cnt.value)
    check(lib().fsx_cg_history(ctx.h, out.ctypes.data_as(C.POINTER(C.c_double)),
                               cnt.value, C.byref(cnt)))
    return out


def ir_cg_solve(A64: CSR, A32: CSR, b: Vector, x: Vector, rtol_outer: float,
                inner: Settings, jacobi: int = 1, max_outer: int = 50) -> IRInfo:
    """jacobi: 0 none, 1 Jacobi-preconditioned inner CG, 2 symmetrically
    scaled inner matrix (A32.set_values_sym_scaled(A64) first)."""
    info = IRInfo()
    check(lib().fsx_ir_cg_solve(A64.ctx.h, A64.h, A32.h, int(jacobi), b.h, x.h,
                                rtol_outer, max_outer, C.byref(inner),
                                C.byref(info)), "fsx_ir_cg_solve")
    return info


def ir_cg_rr_solve(A64: CSR, A32: CSR, b: Vector, x: Vector, rtol: float,
                   inner: Settings, max_segments: int = 100) -> IRInfo:
    """FP32 CG on the symmetrically scaled A32 (A32.set_values_sym_scaled(A64)
    first) with FP64 residual replacement after each segment of inner.rtol
    reduction; x is the initial guess."""
    info = IRInfo()
    check(lib().fsx_ir_cg_rr_solve(A64.ctx.h, A64.h, A32.h, b.h, x.h, rtol,
                                   int(max_segments), C.byref(inner), C.byref(info)),
          "fsx_ir_cg_rr_solve")
    return info


def mm_read(path):
    """Matrix Market file -> (nrows, ncols, symmetric, I, J, V), 0-based COO."""
    nr, nc, lines, cap = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    sym = C.c_int()
    p = str(path).encode()
    check(lib().fsx_mm_read_header(p, C.byref(nr), C.byref(nc), C.byref(lines),
                                   C.byref(sym), C.byref(cap)), "fsx_mm_read_header")
    I = np.empty(max(cap.value, 1), np.int64)
    J = np.empty(max(cap.value, 1), np.int64)
    V = np.empty(max(cap.value, 1))
    cnt = C.c_int64()
    check(lib().fsx_mm_read_coo(p, cap.value, _ptr(I), _ptr(J), _ptr(V), C.byref(cnt)),
          "fsx_mm_read_coo")
    n = cnt.value
    return nr.value, nc.value, bool(sym.value), I[:n], J[:n], V[:n]


def coo_tocsr(nrows: int, I, J, V):
    """Stable-by-row COO -> CSR (int64 indices), duplicates kept."""
    I = np.ascontiguousarray(I, np.int64)
    J = np.ascontiguousarray(J, np.int64)
    V = np.ascontiguousarray(V, np.float64)
    rp = np.empty(nrows + 1, np.int64)
    ci = np.empty(max(I.size, 1), np.int64)
    vv = np.empty(max(I.size, 1))
    check(lib().fsx_coo_tocsr(nrows, I.size, _ptr(I), _ptr(J), _ptr(V), _ptr(rp), _ptr(ci),
                              _ptr(vv)), "fsx_coo_tocsr")
    return rp, ci[:I.size], vv[:I.size]


def csr_from_mm(ctx: "Context", path, itype: int = I32) -> "CSR":
    h = C.c_void_p()
    check(lib().fsx_csr_from_mm(ctx.h, str(path).encode(), itype, C.byref(h)),
          "fsx_csr_from_mm")
    return CSR(ctx, None, None, None, _handle=h)


def be_forward(A: CSR, u0: Vector, steps: int, settings: Settings,
               dinv: Vector | None = None):
    """Backward-Euler trajectory u^{s+1} = A^{-1} u^s; returns (u^steps, stats)."""
    uT = Vector(A.ctx, A.nrows)
    st = BEStats()
    check(lib().fsx_be_forward(A.ctx.h, A.h, dinv.h if dinv is not None else None, u0.h,
                               steps, C.byref(settings), uT.h, C.byref(st)), "fsx_be_forward")
    return uT, st


def be_adjoint(A: CSR, u0: Vector, u_target: Vector, steps: int, ckpt: int,
               settings: Settings, dinv: Vector | None = None, AT: CSR | None = None):
    """Loss |u^steps - u_target|^2 and its gradients by the discrete adjoint.
    Returns (loss, G = dL/dA on A's pattern, u^steps, dL/du^0, stats)."""
    G = Vector(A.ctx, A.nnz)
    uT = Vector(A.ctx, A.nrows)
    gu0 = Vector(A.ctx, A.nrows)
    loss = C.c_double()
    st = BEStats()
    check(lib().fsx_be_adjoint(A.ctx.h, A.h, AT.h if AT is not None else None,
                               dinv.h if dinv is not None else None, u0.h, u_target.h,
                               steps, ckpt, C.byref(settings), C.byref(loss), G.h, uT.h,
                               gu0.h, C.byref(st)), "fsx_be_adjoint")
    return loss.value, G, uT, gu0, st


def be_forward_ir(A: CSR, A32: CSR, u0: Vector, steps: int, settings: Settings,
                  inner: Settings):
    """be_forward with each step solved by mixed-precision refinement on A32
    (S A S in FP32: A32.set_values_sym_scaled(A) first)."""
    uT = Vector(A.ctx, A.nrows)
    st = BEStats()
    check(lib().fsx_be_forward_ir(A.ctx.h, A.h, A32.h, u0.h, steps, C.byref(settings),
                                  C.byref(inner), uT.h, C.byref(st)), "fsx_be_forward_ir")
    return uT, st


def be_adjoint_ir(A: CSR, A32: CSR, u0: Vector, u_target: Vector, steps: int, ckpt: int,
                  settings: Settings, inner: Settings):
    """be_adjoint with every solve by mixed-precision refinement on A32
    (symmetric A).  Returns (loss, G, u^steps, dL/du^0, stats)."""
    G = Vector(A.ctx, A.nnz)
    uT = Vector(A.ctx, A.nrows)
    gu0 = Vector(A.ctx, A.nrows)
    loss = C.c_double()
    st = BEStats()
    check(lib().fsx_be_adjoint_ir(A.ctx.h, A.h, A32.h, u0.h, u_target.h, steps, ckpt,
                                  C.byref(settings), C.byref(inner), C.byref(loss), G.h, uT.h,
                                  gu0.h, C.byref(st)), "fsx_be_adjoint_ir")
    return loss.value, G, uT, gu0, st


def be_forward_rr(A: CSR, A32: CSR, u0: Vector, steps: int, settings: Settings,
                  inner: Settings):
    """be_forward with each step solved by the FP32 CG with FP64 residual
    replacement on A32 (inner.rtol: the segment reduction)."""
    uT = Vector(A.ctx, A.nrows)
    st = BEStats()
    check(lib().fsx_be_forward_rr(A.ctx.h, A.h, A32.h, u0.h, steps, C.byref(settings),
                                  C.byref(inner), uT.h, C.byref(st)), "fsx_be_forward_rr")
    return uT, st


def be_adjoint_rr(A: CSR, A32: CSR, u0: Vector, u_target: Vector, steps: int, ckpt: int,
                  settings: Settings, inner: Settings):
    """be_adjoint with every solve by the FP32 CG with FP64 residual
    replacement on A32 (symmetric A).  Returns (loss, G, u^steps, dL/du^0,
    stats)."""
    G = Vector(A.ctx, A.nnz)
    uT = Vector(A.ctx, A.nrows)
    gu0 = Vector(A.ctx, A.nrows)
    loss = C.c_double()
    st = BEStats()
    check(lib().fsx_be_adjoint_rr(A.ctx.h, A.h, A32.h, u0.h, u_target.h, steps, ckpt,
                                  C.byref(settings), C.byref(inner), C.byref(loss), G.h, uT.h,
                                  gu0.h, C.byref(st)), "fsx_be_adjoint_rr")
    return loss.value, G, uT, gu0, st


def krylov(kind: str, A: CSR, b: Vector, x: Vector, settings: Settings,
           prec: int = 0, nrelax: int = 1, hist_cap: int = 0):
    """Reference-shaped solver (cg, cg_generic, fcg, gmres, bicgstab) over
    device vectors.  Returns (Info, residual history array)."""
    info = Info()
    hist = np.zeros(max(hist_cap, 1))
    hp = hist.ctypes.data_as(C.POINTER(C.c_double))
    L = lib()
    if kind == "fcg":
        rc = L.fsx_krylov_fcg(A.ctx.h, A.h, prec, nrelax, b.h, x.h,
                              C.byref(settings), C.byref(info))
    else:
        fn = getattr(L, f"fsx_krylov_{kind}")
        rc = fn(A.ctx.h, A.h, prec, nrelax, b.h, x.h, C.byref(settings),
                C.byref(info), hp, hist_cap)
    check(rc, f"fsx_krylov_{kind}")
    return info, hist[:min(hist_cap, max(info.iters, 0) or hist_cap)]


class ParCSR:
    """Parallel CSR over `ncolours` colours on one device (fsx_parcsr_*;
    topo::csr colouring and mat::parcsr of flecsolve/topo/csr.hh,
    flecsolve/matrices/parcsr.hh).  A (FP64) is copied."""

    def __init__(self, A: CSR, ncolours: int):
        self.ctx = A.ctx
        self.nrows, self.ncols, self.ncolours = A.nrows, A.ncols, int(ncolours)
        self.h = C.c_void_p()
        check(lib().fsx_parcsr_create(A.ctx.h, A.h, int(ncolours), C.byref(self.h)),
              "fsx_parcsr_create")

    def colour_info(self, c: int) -> dict:
        s = np.zeros(5, np.int64)
        check(lib().fsx_parcsr_colour_info(self.h, int(c), _ptr(s)), "fsx_parcsr_colour_info")
        return dict(rows=int(s[0]), cols=int(s[1]), ghosts=int(s[2]), nnz_diag=int(s[3]),
                    nnz_offd=int(s[4]))

    def download(self, c: int) -> dict:
        """Colour c's blocks in the reference's storage (keys as
        oracle.parcsr.split)."""
        i = self.colour_info(c)
        d = dict(rows=i["rows"], cols=i["cols"],
                 diag_rowptr=np.empty(i["rows"] + 1, np.int64),
                 diag_colind=np.empty(i["nnz_diag"], np.int64),
                 diag_vals=np.empty(i["nnz_diag"]),
                 offd_rowptr=np.empty(i["rows"] + 1, np.int64),
                 offd_colind=np.empty(i["nnz_offd"], np.int64),
                 offd_vals=np.empty(i["nnz_offd"]),
                 colmap=np.empty(i["ghosts"], np.int64))
        check(lib().fsx_parcsr_download(
            self.ctx.h, self.h, int(c), _ptr(d["diag_rowptr"]), _ptr(d["diag_colind"]),
            _ptr(d["diag_vals"]), _ptr(d["offd_rowptr"]), _ptr(d["offd_colind"]),
            _ptr(d["offd_vals"]), _ptr(d["colmap"])), "fsx_parcsr_download")
        return d

    def spmv(self, x: Vector, y: Vector):
        check(lib().fsx_parcsr_spmv(self.ctx.h, self.h, x.h, y.h), "fsx_parcsr_spmv")

    def residual(self, b: Vector, x: Vector, r: Vector):
        check(lib().fsx_parcsr_residual(self.ctx.h, self.h, b.h, x.h, r.h),
              "fsx_parcsr_residual")

    def diag_plan(self) -> tuple[int, int]:
        """SpMV row mapping of the diag blocks' matrix (owned by self)."""
        d = C.c_void_p()
        check(lib().fsx_parcsr_diag(self.h, C.byref(d)), "fsx_parcsr_diag")
        k, l = C.c_int(), C.c_int()
        check(lib().fsx_csr_plan(d, C.byref(k), C.byref(l)))
        return k.value, l.value

    def diag_layout_bytes(self, x_bytes: int = 8) -> int:
        """Compulsory bytes of one product of the diag blocks' matrix on its
        layout (fsx_csr_layout_bytes)."""
        d = C.c_void_p()
        check(lib().fsx_parcsr_diag(self.h, C.byref(d)), "fsx_parcsr_diag")
        b = C.c_uint64()
        check(lib().fsx_csr_layout_bytes(d, x_bytes, C.byref(b)), "fsx_csr_layout_bytes")
        return b.value

    def destroy(self):
        if self.h:
            check(lib().fsx_parcsr_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class HybridGS:
    """Hybrid Gauss-Seidel smoother (fsx_gs_*; op::hybrid_gs of
    flecsolve/solvers/mg/gs.hh).  sweep: 0 forward, 1 backward, 2 symmetric."""

    FORWARD, BACKWARD, SYMMETRIC = 0, 1, 2

    def __init__(self, A: CSR, ncolours: int = 1, sweep: int = 2, nrelax: int = 1):
        self.ctx = A.ctx
        self.A = A  # the smoother reads A's arrays: keep A alive
        self.sweep, self.nrelax = sweep, nrelax
        self.h = C.c_void_p()
        check(lib().fsx_gs_create(A.ctx.h, A.h, int(ncolours), C.byref(self.h)),
              "fsx_gs_create")

    def levels(self):
        f, b = C.c_int64(), C.c_int64()
        check(lib().fsx_gs_levels(self.h, C.byref(f), C.byref(b)), "fsx_gs_levels")
        return f.value, b.value

    def apply(self, b: Vector, x: Vector, sweep: int | None = None,
              nrelax: int | None = None) -> None:
        """Relax x in place (hybrid_gs::apply(b, x))."""
        check(lib().fsx_gs_relax(self.ctx.h, self.h,
                                 self.sweep if sweep is None else int(sweep),
                                 self.nrelax if nrelax is None else int(nrelax), b.h, x.h),
              "fsx_gs_relax")

    def destroy(self):
        if self.h:
            check(lib().fsx_gs_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class AmgSettings(C.Structure):
    """fsx_amg_settings (mg::ua::solver_settings); defaults are the reference's
    option defaults (ua.cc), max_levels (required there) 20."""
    _fields_ = [("max_levels", C.c_int), ("min_coarse", C.c_int64), ("maxiter", C.c_int),
                ("atol", C.c_float), ("beta", C.c_float), ("pairwise_passes", C.c_int),
                ("nrelax", C.c_int), ("cycle", C.c_int), ("kappa", C.c_int),
                ("ktol", C.c_float), ("smoother", C.c_int), ("jacobi_weight", C.c_float)]
    V, RELAXED_W, RELAXED_KAPPA, KAPPA_K = 0, 1, 2, 3
    GAUSS_SEIDEL, JACOBI = 0, 1

    def __init__(self, max_levels=20, min_coarse=5, maxiter=10, atol=1e-6, beta=0.25,
                 pairwise_passes=2, nrelax=2, cycle=0, kappa=1, ktol=0.25, smoother=0,
                 jacobi_weight=0.6666666):
        super().__init__(max_levels, min_coarse, maxiter, atol, beta, pairwise_passes,
                         nrelax, cycle, kappa, ktol, smoother, jacobi_weight)


class AMG:
    """Unsmoothed-aggregation AMG (fsx_amg_*; mg::ua::bound_solver)."""

    def __init__(self, A: CSR, settings: AmgSettings | None = None):
        self.ctx = A.ctx
        self.A = A  # the solver reads A's arrays: keep A alive
        self.settings = settings or AmgSettings()
        self.h = C.c_void_p()
        check(lib().fsx_amg_create(A.ctx.h, A.h, C.byref(self.settings), C.byref(self.h)),
              "fsx_amg_create")

    def levels(self):
        """[(rows, stored entries)] per level."""
        nl = C.c_int()
        check(lib().fsx_amg_levels(self.h, C.byref(nl), None, None, 0))
        rows = np.zeros(nl.value, np.int64)
        nnz = np.zeros(nl.value, np.int64)
        check(lib().fsx_amg_levels(self.h, C.byref(nl), _ptr(rows), _ptr(nnz), nl.value))
        return list(zip(rows.tolist(), nnz.tolist()))

    def aggregates(self, level: int) -> np.ndarray:
        out = np.empty(self.levels()[level][0], np.int64)
        check(lib().fsx_amg_aggregates(self.h, level, _ptr(out)), "fsx_amg_aggregates")
        return out

    def level_matrix(self, level: int) -> CSR:
        h = C.c_void_p()
        check(lib().fsx_amg_level_matrix(self.h, level, C.byref(h)), "fsx_amg_level_matrix")
        M = CSR(self.ctx, None, None, None, _handle=h)
        M._owner = self  # owned by the solver
        return M

    def apply(self, b: Vector, x: Vector):
        """x = the solver applied to b; returns (cycles, |b - A x|)."""
        cy, res = C.c_int(), C.c_double()
        check(lib().fsx_amg_apply(self.ctx.h, self.h, b.h, x.h, C.byref(cy), C.byref(res)),
              "fsx_amg_apply")
        return cy.value, res.value

    def cg(self, b: Vector, x: Vector, settings: Settings, flexible=False, hist_cap=0):
        """CG (or FCG) on the fine matrix with this solver as preconditioner."""
        info = Info()
        hist = np.zeros(max(hist_cap, 1))
        check(lib().fsx_krylov_cg_amg(self.ctx.h, self.A.h, self.h, int(flexible), b.h, x.h,
                                      C.byref(settings), C.byref(info),
                                      _ptr(hist) if hist_cap else None, hist_cap),
              "fsx_krylov_cg_amg")
        return info

    def destroy(self):
        if self.h:
            check(lib().fsx_amg_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """Block-diagonal batch of independent systems (fsx_batch_*)."""

    def __init__(self, A: CSR, sys_rows):
        self.ctx = A.ctx
        self.A = A
        self.rows = np.ascontiguousarray(sys_rows, np.int64)
        self.nsys = self.rows.size - 1
        self.h = C.c_void_p()
        check(lib().fsx_batch_create(A.ctx.h, A.h, self.nsys, _ptr(self.rows),
                                     C.byref(self.h)), "fsx_batch_create")

    def config(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().fsx_batch_config(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_config(self, wg, mapping, lanes=1):
        check(lib().fsx_batch_set_config(self.h, wg, mapping, lanes), "set_config")

    def tune(self, b: Vector, x: Vector, dinv: Vector | None = None, sample=32,
             iters=20) -> float:
        ms = C.c_double()
        check(lib().fsx_batch_tune(self.ctx.h, self.h, dinv.h if dinv else None,
                                   b.h, x.h, sample, iters, C.byref(ms)), "tune")
        return ms.value

    def cg(self, b: Vector, x: Vector, settings: Settings, dinv: Vector | None = None):
        it = np.zeros(self.nsys, np.int32)
        st = np.zeros(self.nsys, np.int32)
        rs = np.zeros(self.nsys, np.float64)
        check(lib().fsx_batched_cg(self.ctx.h, self.h, dinv.h if dinv else None,
                                   b.h, x.h, C.byref(settings), _ptr(it), _ptr(st),
                                   _ptr(rs)), "fsx_batched_cg")
        return it, st, rs

    def ir_cg(self, A32: CSR, b: Vector, x: Vector, rtol_outer: float,
              inner: Settings, jacobi=True, max_outer=20):
        outer = np.zeros(self.nsys, np.int32)
        inn = np.zeros(self.nsys, np.int32)
        st = np.zeros(self.nsys, np.int32)
        rs = np.zeros(self.nsys, np.float64)
        check(lib().fsx_batched_ir_cg(self.ctx.h, self.h, A32.h, int(jacobi), b.h, x.h,
                                      rtol_outer, max_outer, C.byref(inner),
                                      _ptr(outer), _ptr(inn), _ptr(st), _ptr(rs)),
              "fsx_batched_ir_cg")
        return outer, inn, st, rs

    def rr_cg(self, A32: CSR, b: Vector, x: Vector, rtol: float, inner: Settings,
              max_segments=100):
        """Per system: FP32 CG on the symmetrically scaled A32 with FP64
        residual replacement; x is the initial guess.  Returns (segments,
        FP32 iterations, status, true |r|) per system."""
        seg = np.zeros(self.nsys, np.int32)
        inn = np.zeros(self.nsys, np.int32)
        st = np.zeros(self.nsys, np.int32)
        rs = np.zeros(self.nsys, np.float64)
        check(lib().fsx_batched_rr_cg(self.ctx.h, self.h, A32.h, b.h, x.h, rtol,
                                      int(max_segments), C.byref(inner), _ptr(seg),
                                      _ptr(inn), _ptr(st), _ptr(rs)), "fsx_batched_rr_cg")
        return seg, inn, st, rs

    def destroy(self):
        if self.h:
            check(lib().fsx_batch_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
