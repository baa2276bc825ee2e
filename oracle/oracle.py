"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, and nowhere else.  The oracle restates the
reference's serial algorithm (see flecsolve_oracle.cpp for file:line
citations and the list of pinned functions).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

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
    def reason(self):
        return STOP_REASONS[self.status]


class _Csr(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64),
                ("itype", C.c_int), ("rowptr", C.c_void_p), ("colind", C.c_void_p),
                ("vals", C.c_void_p)]


class _Prec(C.Structure):
    _fields_ = [("kind", C.c_int), ("M", C.POINTER(_Csr)), ("omega", C.c_float),
                ("nrelax", C.c_int)]


DIAG_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int64,
                      C.c_double)

_lib = None
_D = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        i64 = C.c_int64
        for nm in ["or_copy"]:
            getattr(L, nm).argtypes = [P, P, i64]
        L.or_set_scalar.argtypes = [C.c_double, P, i64]
        L.or_scale.argtypes = [C.c_double, P, i64]
        L.or_scale_to.argtypes = [C.c_double, P, P, i64]
        for nm in ["or_add", "or_subtract", "or_multiply", "or_divide"]:
            getattr(L, nm).argtypes = [P, P, P, i64]
        L.or_reciprocal.argtypes = [P, P, i64]
        L.or_linear_sum.argtypes = [C.c_double, P, C.c_double, P, P, i64]
        L.or_axpy.argtypes = [C.c_double, P, P, P, i64]
        L.or_axpby.argtypes = [C.c_double, C.c_double, P, P, i64]
        L.or_abs.argtypes = [P, P, i64]
        L.or_add_scalar.argtypes = [P, C.c_double, P, i64]
        for nm in ["or_min", "or_max", "or_l1norm", "or_l2norm", "or_inf_norm"]:
            getattr(L, nm).argtypes = [P, i64]
            getattr(L, nm).restype = C.c_double
        L.or_dot.argtypes = [P, P, i64]
        L.or_dot.restype = C.c_double
        L.or_set_random.argtypes = [P, i64, C.c_uint]
        L.or_spmv.argtypes = [C.POINTER(_Csr), P, P]
        L.or_residual.argtypes = [C.POINTER(_Csr), P, P, P]
        L.or_coo_tocsr.argtypes = [i64, i64, P, P, P, P, P, P]
        L.or_dinv_csr.argtypes = [C.POINTER(_Csr), P, P, P]
        L.or_mm_read.argtypes = [C.c_char_p, C.POINTER(i64), C.POINTER(i64),
                                 C.POINTER(C.c_int), P, P, P, i64]
        L.or_mm_read.restype = i64
        L.or_jacobi_relax.argtypes = [C.POINTER(_Csr), C.c_float, C.c_int, P, P]
        L.or_hybrid_gs.argtypes = [C.POINTER(_Csr), C.c_int, C.c_int, C.c_int, P, P]
        L.or_hybrid_gs.restype = C.c_int
        L.or_amg_create.argtypes = [C.POINTER(_Csr), C.POINTER(AmgSettings)]
        L.or_amg_create.restype = C.c_void_p
        L.or_amg_nlevels.argtypes = [C.c_void_p]
        L.or_amg_nlevels.restype = C.c_int
        L.or_amg_level_rows.argtypes = [C.c_void_p, C.c_int]
        L.or_amg_level_rows.restype = C.c_int64
        L.or_amg_level_nnz.argtypes = [C.c_void_p, C.c_int]
        L.or_amg_level_nnz.restype = C.c_int64
        L.or_amg_level_csr.argtypes = [C.c_void_p, C.c_int, P, P, P]
        L.or_amg_aggregates.argtypes = [C.c_void_p, C.c_int, P]
        L.or_amg_apply.argtypes = [C.c_void_p, P, P, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double)]
        L.or_amg_apply.restype = C.c_int
        L.or_amg_destroy.argtypes = [C.c_void_p]
        L.or_pairwise_agg.argtypes = [C.POINTER(_Csr), C.c_float, C.c_int, P]
        L.or_pairwise_agg.restype = C.c_int64
        solver_args = [C.POINTER(_Csr), C.POINTER(_Prec), P, P, C.POINTER(Settings),
                       DIAG_FN, P, C.POINTER(Info)]
        L.or_cg.argtypes = solver_args + [_D, C.c_int]
        L.or_fcg.argtypes = solver_args
        L.or_gmres.argtypes = solver_args + [_D, C.c_int]
        L.or_bicgstab.argtypes = solver_args + [_D, C.c_int]
        for nm in ["or_cg", "or_fcg", "or_gmres", "or_bicgstab"]:
            getattr(L, nm).restype = C.c_int
        L.or_cg_colours.argtypes = [C.POINTER(_Csr), P, P, P, C.POINTER(Settings), C.c_int,
                                    C.POINTER(Info)]
        L.or_cg_colours.restype = C.c_int
        L.or_lp_powsum_f32.argtypes = [P, i64, C.c_int]
        L.or_lp_powsum_f32.restype = C.c_float
        L.or_mutant_bicgstab_relative_early_exit.argtypes = [C.c_int]
        L.or_mutant_bicgstab_relative_early_exit.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Csr:
    """Holds numpy arrays alive for the oracle's CSR view."""

    def __init__(self, rowptr, colind, vals, ncols=None):
        rowptr = np.ascontiguousarray(rowptr)
        if rowptr.dtype not in (np.int32, np.int64):
            rowptr = rowptr.astype(np.int64)
        idt = rowptr.dtype
        self.rowptr = rowptr
        self.colind = np.ascontiguousarray(colind, dtype=idt)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.nrows = rowptr.size - 1
        self.ncols = self.nrows if ncols is None else ncols
        self.s = _Csr(self.nrows, self.ncols, self.colind.size,
                      0 if idt == np.int32 else 1, _p(self.rowptr),
                      _p(self.colind), _p(self.vals))

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.nrows)
        lib().or_spmv(C.byref(self.s), _p(x), _p(y))
        return y

    def residual(self, b, x):
        b = np.ascontiguousarray(b, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        r = np.empty(self.nrows)
        lib().or_residual(C.byref(self.s), _p(b), _p(x), _p(r))
        return r

    def dinv(self) -> "Csr":
        rp = np.empty(self.nrows + 1, np.int64)
        ci = np.empty(self.nrows, np.int64)
        v = np.empty(self.nrows)
        lib().or_dinv_csr(C.byref(self.s), _p(rp), _p(ci), _p(v))
        return Csr(rp, ci, v)

    def jacobi_relax(self, b, x, omega=np.float32(2 / 3.), nrelax=1):
        x = np.array(x, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        lib().or_jacobi_relax(C.byref(self.s), float(omega), nrelax, _p(b), _p(x))
        return x

    def hybrid_gs(self, b, x, sweep=2, nrelax=1, ncolours=1):
        """op::hybrid_gs relaxation (sweep 0 forward, 1 backward, 2 symmetric)
        over `ncolours` row blocks; returns the relaxed x."""
        x = np.array(x, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        if lib().or_hybrid_gs(C.byref(self.s), int(ncolours), int(sweep), int(nrelax),
                              _p(b), _p(x)) != 0:
            raise ValueError("hybrid_gs: bad colours or sweep")
        return x


class AmgSettings(C.Structure):
    """or_amg_settings (mg::ua::solver_settings)."""
    _fields_ = [("max_levels", C.c_int), ("min_coarse", C.c_int64), ("maxiter", C.c_int),
                ("atol", C.c_float), ("beta", C.c_float), ("pairwise_passes", C.c_int),
                ("nrelax", C.c_int), ("cycle", C.c_int), ("kappa", C.c_int),
                ("ktol", C.c_float), ("smoother", C.c_int), ("jacobi_weight", C.c_float)]

    def __init__(self, max_levels=20, min_coarse=5, maxiter=10, atol=1e-6, beta=0.25,
                 pairwise_passes=2, nrelax=2, cycle=0, kappa=1, ktol=0.25, smoother=0,
                 jacobi_weight=0.6666666):
        super().__init__(max_levels, min_coarse, maxiter, atol, beta, pairwise_passes,
                         nrelax, cycle, kappa, ktol, smoother, jacobi_weight)


class Amg:
    """The oracle's mg::ua::bound_solver (setup and apply, one colour)."""

    def __init__(self, A: "Csr", settings: AmgSettings | None = None):
        self.A = A
        self.settings = settings or AmgSettings()
        self.h = lib().or_amg_create(C.byref(A.s), C.byref(self.settings))
        if not self.h:
            raise ValueError("or_amg_create: bad settings or matrix")

    def nlevels(self):
        return lib().or_amg_nlevels(self.h)

    def level(self, l):
        n = lib().or_amg_level_rows(self.h, l)
        nz = lib().or_amg_level_nnz(self.h, l)
        rp, ci, v = np.empty(n + 1, np.int64), np.empty(nz, np.int64), np.empty(nz)
        lib().or_amg_level_csr(self.h, l, _p(rp), _p(ci), _p(v))
        return rp, ci, v

    def aggregates(self, l):
        out = np.empty(lib().or_amg_level_rows(self.h, l), np.int64)
        lib().or_amg_aggregates(self.h, l, _p(out))
        return out

    def apply(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        cy, res = C.c_int(), C.c_double()
        lib().or_amg_apply(self.h, _p(b), _p(x), C.byref(cy), C.byref(res))
        return x, cy.value, res.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_amg_destroy(self.h)
            self.h = None


def pairwise_agg(A: "Csr", beta=0.25, checkdd=True):
    """One pairwise aggregation pass: aggregate id per row (-1: none), count."""
    out = np.empty(A.s.nrows, np.int64)
    k = lib().or_pairwise_agg(C.byref(A.s), float(beta), int(checkdd), _p(out))
    return out, k


def prec(kind: int = 0, M: Csr | None = None, omega=np.float32(2 / 3.),
         nrelax: int = 1):
    """0 identity, 1 CSR operator M (e.g. Csr.dinv()), 2 Jacobi relaxation on M."""
    p = _Prec(kind, C.pointer(M.s) if M is not None else None, float(omega), nrelax)
    p._keep = M
    return p


def _solve(name, A: Csr, b, x0, settings: Settings, P=None, diag=None,
           hist_cap=0):
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, dtype=np.float64, copy=True)
    info = Info()
    if diag is None:
        fn = DIAG_FN(0)
    else:
        def _cb(_u, xp, n, res):
            return int(bool(diag(np.ctypeslib.as_array(xp, shape=(n,)), res)))
        fn = DIAG_FN(_cb)
    Pp = C.byref(P) if P is not None else None
    args = [C.byref(A.s), Pp, _p(b), _p(x), C.byref(settings), fn, None,
            C.byref(info)]
    hist = np.zeros(max(hist_cap, 1))
    if name != "or_fcg":
        args += [hist.ctypes.data_as(_D), hist_cap]
    rc = getattr(lib(), name)(*args)
    return x, info, rc, hist[:hist_cap]


def cg_colours(A: Csr, dinv, b, x0, settings: Settings, ncolours: int):
    """CG + Jacobi over `ncolours` row blocks on as many threads, block
    partials summed in block order (the reference's MPI backend with one
    colour per rank).  CPU baseline only."""
    b = np.ascontiguousarray(b, np.float64)
    d = np.ascontiguousarray(dinv, np.float64)
    x = np.array(x0, dtype=np.float64, copy=True)
    info = Info()
    rc = lib().or_cg_colours(C.byref(A.s), _p(d), _p(b), _p(x), C.byref(settings),
                             int(ncolours), C.byref(info))
    if rc != 0:
        raise ValueError("ncolours must be >= 1")
    return x, info


def cg(A, b, x0, settings, P=None, diag=None, hist_cap=0):
    x, info, rc, h = _solve("or_cg", A, b, x0, settings, P, diag, hist_cap)
    return x, info, h


def fcg(A, b, x0, settings, P=None, diag=None):
    x, info, rc, _ = _solve("or_fcg", A, b, x0, settings, P, diag)
    return x, info


def gmres(A, b, x0, settings, P=None, diag=None, hist_cap=0):
    x, info, rc, h = _solve("or_gmres", A, b, x0, settings, P, diag, hist_cap)
    if rc != 0:
        raise ValueError("gmres settings rejected (krylov dimension bound)")
    return x, info, h


def bicgstab(A, b, x0, settings, P=None, diag=None, hist_cap=0):
    x, info, rc, h = _solve("or_bicgstab", A, b, x0, settings, P, diag, hist_cap)
    return x, info, h


class mutant_relative_early_exit:
    """Context manager planting a mistake in the oracle's BiCGSTAB (tests
    only): inside it, the early exit compares |s| with rtol * |b| instead of
    the absolute rtol of bicgstab.hh:136-142."""

    def __enter__(self):
        lib().or_mutant_bicgstab_relative_early_exit(1)
        return self

    def __exit__(self, *exc):
        lib().or_mutant_bicgstab_relative_early_exit(0)
        return False


# ---- BLAS-1 ----------------------------------------------------------------

def _a(x):
    return np.ascontiguousarray(x, np.float64)


def dot(x, y):
    x, y = _a(x), _a(y)
    return lib().or_dot(_p(x), _p(y), x.size)


def l2norm(x):
    x = _a(x)
    return lib().or_l2norm(_p(x), x.size)


def l1norm(x):
    x = _a(x)
    return lib().or_l1norm(_p(x), x.size)


def inf_norm(x):
    x = _a(x)
    return lib().or_inf_norm(_p(x), x.size)


def vmin(x):
    x = _a(x)
    return lib().or_min(_p(x), x.size)


def vmax(x):
    x = _a(x)
    return lib().or_max(_p(x), x.size)


def lp_powsum_f32(x, p):
    """FP32 lp_norm_local: float terms folded in float, index order."""
    x = np.ascontiguousarray(x, np.float32)
    return float(lib().or_lp_powsum_f32(x.ctypes.data_as(C.c_void_p), x.size, int(p)))


def set_random(n, seed):
    x = np.empty(n)
    lib().or_set_random(_p(x), n, seed)
    return x


def binop(name, x, y):
    x, y = _a(x), _a(y)
    z = np.empty_like(x)
    getattr(lib(), "or_" + name)(_p(x), _p(y), _p(z), x.size)
    return z


def axpy(a, x, y):
    x, y = _a(x), _a(y)
    z = np.empty_like(x)
    lib().or_axpy(a, _p(x), _p(y), _p(z), x.size)
    return z


def axpby(a, b, x, z):
    x, z = _a(x), np.array(z, np.float64)
    lib().or_axpby(a, b, _p(x), _p(z), x.size)
    return z


def linear_sum(a, x, b, y):
    x, y = _a(x), _a(y)
    z = np.empty_like(x)
    lib().or_linear_sum(a, _p(x), b, _p(y), _p(z), x.size)
    return z


def unop(name, x, *scal):
    x = _a(x)
    y = np.empty_like(x)
    if name == "scale_to":
        lib().or_scale_to(scal[0], _p(x), _p(y), x.size)
    elif name == "add_scalar":
        lib().or_add_scalar(_p(x), scal[0], _p(y), x.size)
    else:
        getattr(lib(), "or_" + name)(_p(x), _p(y), x.size)
    return y


def coo_tocsr(nrows, I, J, V):
    I = np.ascontiguousarray(I, np.int64)
    J = np.ascontiguousarray(J, np.int64)
    V = _a(V)
    rp = np.empty(nrows + 1, np.int64)
    ci = np.empty(I.size, np.int64)
    vv = np.empty(I.size)
    lib().or_coo_tocsr(nrows, I.size, _p(I), _p(J), _p(V), _p(rp), _p(ci), _p(vv))
    return rp, ci, vv


def mm_read(path):
    nr, nc, sym = C.c_int64(), C.c_int64(), C.c_int()
    cap = lib().or_mm_read(str(path).encode(), C.byref(nr), C.byref(nc),
                           C.byref(sym), None, None, None, 0)
    if cap < 0:
        raise IOError(path)
    I = np.empty(cap, np.int64)
    J = np.empty(cap, np.int64)
    V = np.empty(cap)
    cnt = lib().or_mm_read(str(path).encode(), C.byref(nr), C.byref(nc), C.byref(sym),
                           _p(I), _p(J), _p(V), cap)
    if cnt < 0:
        raise IOError(path)
    return nr.value, nc.value, bool(sym.value), I[:cnt], J[:cnt], V[:cnt]


# ---- backward-Euler diffusion and its adjoint --------------------------------
# The forward step restates the reference's BE integrator
# (time-integrators/bdf.hh:86-94, 269-272, 292-296 with operator_adapter.hh:
# 29-35): solve (I + gamma F) u^{s+1} = u^s from the guess u^s.  The adjoint
# has no reference counterpart; it is pinned by central finite differences of
# the forward loss (tests/test_oracle_adjoint.py) and, for the coefficient
# chain, by finite differences of a linear functional of the assembled values.

def be_forward(A: Csr, u0, steps, settings, jacobi=True):
    """States u^0..u^steps of u^{s+1} = A^{-1} u^s (CG, Jacobi if asked)."""
    P = prec(1, A.dinv()) if jacobi else None
    states = [np.array(u0, np.float64)]
    iters = 0
    for s in range(steps):
        x, info, _ = cg(A, states[-1], states[-1], settings, P=P)
        if info.reason not in ("converged_rtol", "converged_atol"):
            raise RuntimeError(f"oracle BE step {s}: {info.reason}")
        iters += info.iters
        states.append(x)
    return states, iters


def sampled_outer(rowptr, colind, lam, x, scale, out):
    """out[e] += (scale * lam[row e]) * x[col e] over the stored entries."""
    rows = np.repeat(np.arange(rowptr.size - 1), np.diff(rowptr))
    out += (scale * lam[rows]) * x[colind]
    return out


def be_adjoint(A: Csr, u0, u_target, steps, settings, jacobi=True):
    """Loss |u^N - u_target|^2, G = dL/dA on A's pattern and dL/du^0, keeping
    every state (no checkpoints).  A is symmetric here, so A^{-T} = A^{-1}."""
    states, _ = be_forward(A, u0, steps, settings, jacobi)
    P = prec(1, A.dinv()) if jacobi else None
    d = states[-1] - u_target
    loss = float(np.dot(d, d))
    lam = 2.0 * d
    G = np.zeros(A.colind.size)
    for s in range(steps - 1, -1, -1):
        mu, info, _ = cg(A, lam, lam, settings, P=P)
        if info.reason not in ("converged_rtol", "converged_atol"):
            raise RuntimeError(f"oracle adjoint step {s}: {info.reason}")
        sampled_outer(A.rowptr, A.colind, mu, states[s + 1], -1.0, G)
        lam = mu
    return loss, G, lam, states[-1]


def _entry_pos(rowptr, colind, rows, cols, width=7):
    """Position of column cols[q] in row rows[q] (rows of at most `width`)."""
    pos = np.full(rows.size, -1, np.int64)
    lo, hi = rowptr[rows], rowptr[rows + 1]
    for j in range(width):
        e = lo + j
        ok = (e < hi) & (pos < 0)
        e = np.where(ok, e, 0)
        hit = ok & (colind[e] == cols)
        pos[hit] = e[hit]
    assert (pos >= 0).all(), "entry not in the pattern"
    return pos


def diffusion3d_coeff_grad(n, rowptr, colind, G, k, scale):
    """dL/dlog k on the (n+2)^3 lattice from G = dL/dA for A = shift I +
    scale L_k (7-point, k_f = (k_a + k_b)/2): face by face, g_f = ((G_ll +
    G_hh) - G_lh) - G_hl over the face's lower/upper nodes (Dirichlet nodes
    contribute no term), each node summing (scale/2) g_f over its faces in the
    order -z, -y, -x, +x, +y, +z, times k at the node."""
    L = n + 2
    rowptr = np.asarray(rowptr, np.int64)
    colind = np.asarray(colind, np.int64)
    kk, jj, ii = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    inside = (ii >= 1) & (ii <= n) & (jj >= 1) & (jj <= n) & (kk >= 1) & (kk <= n)
    R = np.where(inside, (ii - 1) + n * ((jj - 1) + n * (kk - 1)), -1)
    gdiag = np.zeros((L, L, L))
    rows = R[inside]
    gdiag[inside] = G[_entry_pos(rowptr, colind, rows, rows)]
    hs = 0.5 * scale

    def face_terms(axis):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = slice(0, L - 1)
        hi[axis] = slice(1, L)
        lo, hi = tuple(lo), tuple(hi)
        lin, hin = inside[lo], inside[hi]
        both = lin & hin
        glh = np.zeros(lin.shape)
        ghl = np.zeros(lin.shape)
        rl, rh = R[lo][both], R[hi][both]
        glh[both] = G[_entry_pos(rowptr, colind, rl, rh)]
        ghl[both] = G[_entry_pos(rowptr, colind, rh, rl)]
        gll = np.where(lin, gdiag[lo], 0.0)
        ghh = np.where(hin, gdiag[hi], 0.0)
        t = hs * (((gll + ghh) - glh) - ghl)
        return np.where(lin | hin, t, 0.0)

    tz, ty, tx = face_terms(0), face_terms(1), face_terms(2)
    acc = np.zeros((L, L, L))
    acc[1:, :, :] += tz    # -z face of each node
    acc[:, 1:, :] += ty    # -y
    acc[:, :, 1:] += tx    # -x
    acc[:, :, :-1] += tx   # +x
    acc[:, :-1, :] += ty   # +y
    acc[:-1, :, :] += tz   # +z
    return (np.asarray(k).reshape(L, L, L) * acc).ravel()
