"""Device BLAS-1 and reductions against the oracle (gpu)."""
import numpy as np
import pytest

from flecsolve_amd import fsx
from oracle import oracle as O
from tests import vector_kat

pytestmark = pytest.mark.gpu


class DeviceOps:
    def __init__(self, ctx):
        self.ctx = ctx
        self.L = fsx.lib()

    def new(self, v):
        return self.ctx.vec(np.asarray(v, np.float64))

    def get(self, h):
        return h.get()

    def _c(self, rc):
        fsx.check(rc)

    def add(self, z, x, y): self._c(self.L.fsx_add(self.ctx.h, x.h, y.h, z.h))
    def subtract(self, z, x, y): self._c(self.L.fsx_subtract(self.ctx.h, x.h, y.h, z.h))
    def multiply(self, z, x, y): self._c(self.L.fsx_multiply(self.ctx.h, x.h, y.h, z.h))
    def divide(self, z, x, y): self._c(self.L.fsx_divide(self.ctx.h, x.h, y.h, z.h))
    def add_scalar(self, y, x, a): self._c(self.L.fsx_add_scalar(self.ctx.h, x.h, a, y.h))
    def scale_to(self, y, a, x): self._c(self.L.fsx_scale_to(self.ctx.h, a, x.h, y.h))
    def reciprocal(self, y, x): self._c(self.L.fsx_reciprocal(self.ctx.h, x.h, y.h))
    def linear_sum(self, z, a, x, b, y): self._c(self.L.fsx_linear_sum(self.ctx.h, a, x.h, b, y.h, z.h))
    def axpy(self, z, a, x, y): self._c(self.L.fsx_axpy(self.ctx.h, a, x.h, y.h, z.h))
    def copy(self, d, s): self._c(self.L.fsx_copy(self.ctx.h, s.h, d.h))
    def axpby(self, z, a, b, x): self._c(self.L.fsx_axpby(self.ctx.h, a, b, x.h, z.h))
    def abs(self, y, x): self._c(self.L.fsx_abs(self.ctx.h, x.h, y.h))
    def set_scalar(self, x, a): self._c(self.L.fsx_set_scalar(self.ctx.h, a, x.h))
    def min(self, x): return fsx.vmin(x)
    def max(self, x): return fsx.vmax(x)
    def dot(self, x, y): return fsx.dot(x, y)
    def l1norm(self, x): return fsx.nrm1(x)
    def l2norm(self, x): return fsx.nrm2(x)


def test_flecsi_vector_closed_forms(ctx):
    vector_kat.run(DeviceOps(ctx))


@pytest.mark.parametrize("n", [1, 255, 256, 257, 100_003, 2_000_000])
def test_elementwise_bit_exact(ctx, n):
    rng = np.random.default_rng(n)
    xa, ya = rng.standard_normal(n), rng.standard_normal(n) + 3.0
    x, y, z = ctx.vec(xa), ctx.vec(ya), ctx.zeros(n)
    D = DeviceOps(ctx)
    for name in ["add", "subtract", "multiply", "divide"]:
        getattr(D, name)(z, x, y)
        np.testing.assert_array_equal(z.get(), O.binop(name, xa, ya), err_msg=name)
    D.axpy(z, -1.75, x, y)
    np.testing.assert_array_equal(z.get(), O.axpy(-1.75, xa, ya))
    D.linear_sum(z, 0.3, x, -2.0, y)
    np.testing.assert_array_equal(z.get(), O.linear_sum(0.3, xa, -2.0, ya))
    D.copy(z, y)
    D.axpby(z, 1.5, -0.25, x)
    np.testing.assert_array_equal(z.get(), O.axpby(1.5, -0.25, xa, ya))
    D.reciprocal(z, y)
    np.testing.assert_array_equal(z.get(), O.unop("reciprocal", ya))
    D.abs(z, x)
    np.testing.assert_array_equal(z.get(), np.abs(xa))
    D.add_scalar(z, x, 0.125)
    np.testing.assert_array_equal(z.get(), xa + 0.125)
    D.scale_to(z, 3.0, x)
    np.testing.assert_array_equal(z.get(), 3.0 * xa)


def test_aliasing_forms(ctx):
    # the solvers call axpy with the output aliasing either input
    # (cg.hh:108-109, 130); subtract(b, r) into r is the residual form
    n = 5000
    rng = np.random.default_rng(2)
    xa, ya = rng.standard_normal(n), rng.standard_normal(n)
    L = fsx.lib()
    x, y = ctx.vec(xa), ctx.vec(ya)
    fsx.check(L.fsx_axpy(ctx.h, 2.0, x.h, y.h, x.h))     # x = 2x + y
    np.testing.assert_array_equal(x.get(), 2.0 * xa + ya)
    x.upload(xa)
    fsx.check(L.fsx_axpy(ctx.h, 2.0, x.h, y.h, y.h))     # y = 2x + y
    np.testing.assert_array_equal(y.get(), 2.0 * xa + ya)
    y.upload(ya)
    fsx.check(L.fsx_subtract(ctx.h, y.h, x.h, x.h))      # x = y - x
    np.testing.assert_array_equal(x.get(), ya - xa)
    # scale(alpha, x, y) with x == y is a contract violation (topo_view.hh:72)
    rc = L.fsx_scale_to(ctx.h, 2.0, x.h, x.h)
    assert rc == 1 and b"cannot be the same" in L.fsx_last_error()


@pytest.mark.parametrize("n", [1, 1000, 777_777, 16_777_216])
def test_reductions(ctx, n):
    rng = np.random.default_rng(n + 1)
    xa, ya = rng.standard_normal(n), rng.standard_normal(n)
    x, y = ctx.vec(xa), ctx.vec(ya)
    # serial and tree summation orders differ by O(sqrt(n) eps) of sum |terms|
    tol = 16 * np.sqrt(n) * np.finfo(np.float64).eps
    scale = np.abs(xa * ya).sum()
    assert abs(fsx.dot(x, y) - O.dot(xa, ya)) <= tol * scale + 1e-300
    assert abs(fsx.nrm2(x) - O.l2norm(xa)) <= tol * O.l2norm(xa)
    assert abs(fsx.nrm1(x) - O.l1norm(xa)) <= tol * O.l1norm(xa)
    assert fsx.nrminf(x) == O.inf_norm(xa)
    assert fsx.vmin(x) == O.vmin(xa)
    assert fsx.vmax(x) == O.vmax(xa)
    # deterministic: same bits on repeat
    assert fsx.dot(x, y) == fsx.dot(x, y)


def test_reduction_edge_cases(ctx):
    z = ctx.zeros(10)
    assert fsx.nrminf(z) == np.finfo(np.float64).tiny  # seq.hh:251 quirk
    assert fsx.nrm2(z) == 0.0
    e = ctx.zeros(0)
    assert fsx.vmin(e) == np.finfo(np.float64).max
    assert fsx.dot(e, e) == 0.0


@pytest.mark.parametrize("n", [1, 1000, 300_001])
def test_lp_norm_general_p(ctx, n):
    """lp_norm<p> for p >= 3 as the FleCSI vector path computes it
    (vectors/operations/topo_tasks.hh:216-226): (sum x^p)^(1/p) with no
    absolute value; p = 1, 2 are the l1 / l2 norms.  Closed forms: x = c
    (constant) gives c n^(1/p); x = [-1, 2] at p = 3 gives 7^(1/3) where the
    absolute-value form would give 9^(1/3)."""
    rng = np.random.default_rng(n)
    x = rng.random(n) + 0.5
    xd = ctx.vec(x)
    assert fsx.nrmp(xd, 1) == fsx.nrm1(xd)
    assert fsx.nrmp(xd, 2) == fsx.nrm2(xd)
    x32 = x.astype(np.float32)
    for p in (3, 4, 7):
        want = np.sum(x ** p) ** (1.0 / p)
        assert abs(fsx.nrmp(xd, p) - want) <= 1e-13 * want, p
        # FP32: float terms folded in float (the oracle's serial fold); two
        # float folds of positive terms in different orders differ by at most
        # 2 (n - 1) 2^-24 of the sum, 1/p of that after the root
        s32 = O.lp_powsum_f32(x32, p)
        got = fsx.nrmp(ctx.vec(x32), p)
        assert abs(got - s32 ** (1.0 / p)) <= (2 * max(n - 1, 1) * 2.0 ** -24 / p + 1e-15) \
            * s32 ** (1.0 / p), p
    # float fold pinned exactly: 16 terms (1 + 2^-10)^3 rounded to float sum
    # without rounding in any order; the double fold gives other bits
    e = np.full(16, 1 + 2.0 ** -10, np.float32)
    t = 1 + 3 * 2.0 ** -10 + 3 * 2.0 ** -20
    assert fsx.nrmp(ctx.vec(e), 3) == (16 * t) ** (1.0 / 3)
    assert (16 * t) ** (1.0 / 3) != (16 * (1 + 2.0 ** -10) ** 3) ** (1.0 / 3)
    c = ctx.vec(np.full(n, 0.75))
    assert abs(fsx.nrmp(c, 5) - 0.75 * n ** 0.2) <= 1e-13
    assert abs(fsx.nrmp(ctx.vec(np.array([-1.0, 2.0])), 3) - 7 ** (1 / 3)) <= 1e-15
    assert np.isnan(fsx.nrmp(ctx.vec(np.array([-2.0, 1.0])), 3))
    assert fsx.nrmp(ctx.zeros(0), 3) == 0.0
    with pytest.raises(fsx.FsxError):
        fsx.nrmp(xd, 0)


def test_set_random_matches_reference_stream(ctx):
    v = ctx.zeros(1001)
    fsx.check(fsx.lib().fsx_set_random(ctx.h, v.h, 7))
    np.testing.assert_array_equal(v.get(), O.set_random(1001, 7))


def test_fp32_and_convert(ctx):
    n = 12345
    a = np.random.default_rng(3).standard_normal(n)
    v64 = ctx.vec(a)
    v32 = ctx.zeros(n, fsx.F32)
    fsx.check(fsx.lib().fsx_convert(ctx.h, v64.h, v32.h, 0.5))
    np.testing.assert_array_equal(v32.get(), (0.5 * a).astype(np.float32))
    assert abs(fsx.dot(v32, v32) - float(np.dot(v32.get().astype(np.float64),
                                                v32.get().astype(np.float64)))) < 1e-9 * n


def test_vector_identity_and_refcount(ctx):
    a = ctx.zeros(10)
    b = ctx.zeros(10)
    assert a.id() != b.id()
    fsx.check(fsx.lib().fsx_vec_retain(a.h))
    fsx.check(fsx.lib().fsx_vec_release(a.h))
    assert a.get().shape == (10,)


def test_reduce_async_matches_sync(ctx):
    import ctypes as C
    n = 300_001
    x = ctx.vec(O.set_random(n, 3) - 0.5)
    y = ctx.vec(O.set_random(n, 4) - 0.5)
    out = ctx.zeros(8)
    base = C.c_void_p()
    fsx.check(fsx.lib().fsx_vec_device_ptr(out.h, C.byref(base)))
    for k in range(6):
        fsx.check(fsx.lib().fsx_reduce_async(ctx.h, k, x.h, y.h, C.c_void_p(base.value + 8 * k)),
                  "reduce_async")
    got = out.get()
    assert got[0] == fsx.dot(x, y)
    assert got[1] == fsx.nrm1(x)
    assert np.sqrt(got[2]) == fsx.nrm2(x)
    assert got[3] == fsx.nrminf(x)
    xh = x.get()
    assert got[4] == xh.min() and got[5] == xh.max()
    with pytest.raises(fsx.FsxError):
        fsx.check(fsx.lib().fsx_reduce_async(ctx.h, 0, x.h, None, C.c_void_p(base.value)),
                  "dot without y")


@pytest.mark.parametrize("jac", [True, False])
def test_cg_update(ctx, jac):
    import ctypes as C
    n = 100_003
    rng = np.random.default_rng(8)
    p, w, x, r = (rng.standard_normal(n) for _ in range(4))
    dinv = rng.uniform(0.5, 2.0, n)
    alpha = 0.37
    pv, wv, xv, rv = ctx.vec(p), ctx.vec(w), ctx.vec(x), ctx.vec(r)
    dv, zv = ctx.vec(dinv), ctx.zeros(n)
    rr, rz = C.c_double(), C.c_double()
    fsx.check(fsx.lib().fsx_cg_update(ctx.h, alpha, pv.h, wv.h, dv.h if jac else None, xv.h,
                                      rv.h, zv.h if jac else None, C.byref(rr), C.byref(rz)),
              "cg_update")
    xe = O.axpy(alpha, p, x)
    re = O.axpy(-alpha, w, r)
    np.testing.assert_array_equal(xv.get(), xe)
    np.testing.assert_array_equal(rv.get(), re)
    assert abs(rr.value - O.dot(re, re)) <= 1e-12 * O.dot(re, re)
    if jac:
        ze = 0.0 + dinv * re
        np.testing.assert_array_equal(zv.get(), ze)
        assert abs(rz.value - O.dot(re, ze)) <= 1e-12 * np.abs(re * ze).sum()
    else:
        assert rz.value == rr.value


@pytest.mark.parametrize("dtype", [fsx.F64, fsx.F32])
def test_element_access(ctx, dtype):
    """vec::core::operator[] / Ops::retreive (vectors/core.hh:338-340,
    vectors/seq.hh:271-280) on device vectors: reads see queued work,
    writes land in place, out-of-range indices are errors."""
    n = 1000
    x = np.arange(n, dtype=np.float64) * 0.5
    v = ctx.vec(x if dtype == fsx.F64 else x.astype(np.float32))
    fsx.check(fsx.lib().fsx_scale(ctx.h, 2.0, v.h))   # queued, then read
    assert v[0] == 0.0 and v[7] == 7.0 and v[n - 1] == n - 1.0
    v[3] = -1.25
    v[n - 1] = 0.1
    want = x * 2
    want[3] = -1.25
    want[n - 1] = 0.1 if dtype == fsx.F64 else float(np.float32(0.1))
    np.testing.assert_array_equal(v.get(), want.astype(v.np_dtype))
    assert v[n - 1] == want[n - 1]
    with pytest.raises(fsx.FsxError):
        v[n]
    with pytest.raises(fsx.FsxError):
        v[n] = 1.0
