"""Device CSR SpMV and fused kernels against the oracle (gpu).

Bar: structure round trips bit-exact; SpMV on the LDS-staged row-block,
sliced, slice-offset and symmetric stencil layouts bit-identical to the
reference's serial CSR loop (each sums every row serially in CSR order); the
vector-per-row mapping within 1e-13 relative (tree order inside a row).
"""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from flecsolve_amd import fsx
from flecsolve_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def ragged(n, seed, maxlen=300, itype=np.int32):
    """Rows of very different lengths (including empty rows)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, n)
    lens[rng.integers(0, n, max(1, n // 50))] = rng.integers(100, maxlen, max(1, n // 50))
    lens[::97] = 0
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(itype)
    ci = rng.integers(0, n, rp[-1]).astype(itype)
    v = rng.standard_normal(rp[-1])
    return rp, ci, v


def check_spmv(ctx, rp, ci, v, exact, ncols=None, vtype=None):
    A = ctx.csr(rp, ci, v, ncols=ncols, vtype=vtype)
    ncols = A.ncols
    x = np.random.default_rng(9).standard_normal(ncols)
    xd, yd = ctx.vec(x), ctx.zeros(A.nrows)
    A.spmv(xd, yd)
    vals64 = np.asarray(v, np.float64) if vtype != fsx.F32 else np.asarray(v, np.float32).astype(np.float64)
    ref = O.Csr(rp, ci, vals64, ncols=ncols).spmv(x)
    y = yd.get()
    if exact:
        np.testing.assert_array_equal(y, ref)
    else:
        assert rel(y, ref) <= 1e-13
    return A


def test_structure_round_trip(ctx):
    rp, ci, v = P.poisson2d(64)
    A = ctx.csr(rp, ci, v)
    r2, c2, v2 = A.download()
    np.testing.assert_array_equal(r2, rp)
    np.testing.assert_array_equal(c2, ci)
    np.testing.assert_array_equal(v2, v)
    assert A.plan() == (4, 0)  # symmetric stencil: one offset list, symmetric values


@pytest.mark.parametrize("n", [1, 7, 64])
def test_poisson2d_bit_exact(ctx, n):
    rp, ci, v = P.poisson2d(n)
    check_spmv(ctx, rp, ci, v, exact=True)


@pytest.mark.parametrize("n", [5, 32, 96])
def test_diffusion3d_bit_exact(ctx, n):
    k = P.lognormal_field(n, 1.0, 8.0, 2001)
    rp, ci, v = P.diffusion3d(n, k)
    check_spmv(ctx, rp, ci, v, exact=True)


def test_convdiff_bit_exact(ctx):
    rp, ci, v = P.convdiff3d(40)
    check_spmv(ctx, rp, ci, v, exact=True)


@pytest.mark.parametrize("n", [3, 33, 70])
def test_every_plan_on_stencil(ctx, n):
    # plans 0 (LDS-staged), 2 (sliced), 3 (slice offsets) and 4 (symmetric
    # stencil) sum rows serially: bit-exact; plan 1 (lanes per row) sums in a
    # tree: rounding
    k = P.lognormal_field(n, 1.0, 8.0, 11)
    rp, ci, v = P.diffusion3d(n, k)
    A = ctx.csr(rp, ci, v)
    assert A.plan() == ((4, 0) if n > 3 else (0, 0))  # 27 rows: too much padding
    x = np.random.default_rng(3).standard_normal(n ** 3)
    ref = O.Csr(rp, ci, v).spmv(x)
    xd, yd = ctx.vec(x), ctx.zeros(n ** 3)
    for plan in [(0, 0), (2, 0), (3, 0), (4, 0), (1, 2), (1, 8), (1, 64)]:
        A.set_plan(*plan)
        A.spmv(xd, yd)
        if plan[0] != 1:
            np.testing.assert_array_equal(yd.get(), ref, err_msg=str(plan))
        else:
            assert rel(yd.get(), ref) <= 1e-13, plan
    # fused residual and SpMV+dot on the slice layouts
    b = np.random.default_rng(4).standard_normal(n ** 3)
    r = ctx.zeros(n ** 3)
    bd = ctx.vec(b)
    for plan in (2, 3, 4):
        A.set_plan(plan)
        A.residual(bd, xd, r)
        np.testing.assert_array_equal(r.get(), b - ref)
        pw = fsx.spmv_dot(A, xd, yd)
        np.testing.assert_array_equal(yd.get(), ref)
        assert abs(pw - O.dot(ref, x)) <= 1e-12 * np.abs(ref * x).sum()
    # FP32 copies share the index arrays and start on slice offsets; both
    # layouts sum each row serially in FP32, so their results agree bit for
    # bit
    x32 = ctx.vec(x.astype(np.float32))
    ref32 = O.Csr(rp, ci, v.astype(np.float32).astype(np.float64)).spmv(
        x.astype(np.float32).astype(np.float64))
    ys = []
    for plan in (3, 4):
        A.set_plan(plan)
        A32 = A.convert(fsx.F32)
        assert A32.plan() == (3, 0)
        A32.set_plan(plan)
        y32 = ctx.zeros(n ** 3, fsx.F32)
        A32.spmv(x32, y32)
        assert rel(y32.get().astype(np.float64), ref32) <= 1e-6
        ys.append(y32.get())
    np.testing.assert_array_equal(ys[0], ys[1])


def test_stencil_layout_needs_symmetry(ctx):
    """The symmetric stencil layout reads each entry below the diagonal
    from its mirror above, so it serves only bitwise symmetric values and
    structures: asymmetric values at creation, a signed zero mirrored by +0,
    a missing mirror entry, values set asymmetric later (the products fall
    back to the slice-offset layout), and shift-scaled copies; every product
    bit-exact against the oracle."""
    n = 18
    k = P.lognormal_field(n, 1.0, 8.0, 21)
    rp, ci, v = P.diffusion3d(n, k)
    N = n ** 3
    x = np.random.default_rng(5).standard_normal(N)
    xd, yd = ctx.vec(x), ctx.zeros(N)
    rows = np.repeat(np.arange(N), np.diff(rp))
    lower = np.flatnonzero(ci < rows)

    def exact(A, vals):
        A.spmv(xd, yd)
        np.testing.assert_array_equal(yd.get(), O.Csr(rp, ci, vals).spmv(x))

    A = ctx.csr(rp, ci, v)
    assert A.plan() == (4, 0)
    exact(A, v)
    # int64 indices: the same layout
    A64i = ctx.csr(rp.astype(np.int64), ci.astype(np.int64), v)
    assert A64i.plan() == (4, 0)
    exact(A64i, v)
    # an FP32 matrix starts on slice offsets; on request the stencil layout,
    # the same bits
    v32 = v.astype(np.float32)
    F = ctx.csr(rp, ci, v32)
    assert F.plan() == (3, 0)
    x32, y32 = ctx.vec(x.astype(np.float32)), ctx.zeros(N, fsx.F32)
    F.spmv(x32, y32)
    y3 = y32.get()
    F.set_plan(4)
    F.spmv(x32, y32)
    np.testing.assert_array_equal(y32.get(), y3)
    # one entry below the diagonal off by one ulp
    va = v.copy()
    va[lower[100]] = np.nextafter(va[lower[100]], 0.0)
    B = ctx.csr(rp, ci, va)
    assert B.plan() == (3, 0)
    exact(B, va)
    with pytest.raises(fsx.FsxError, match="not symmetric"):
        B.set_plan(4)
    # -0.0 below the diagonal, +0.0 above: equal values, different bits
    vz = v.copy()
    e = lower[7]
    i, j = rows[e], ci[e]
    m = rp[j] + np.flatnonzero(ci[rp[j]:rp[j + 1]] == i)[0]
    vz[e], vz[m] = -0.0, 0.0
    Z = ctx.csr(rp, ci, vz)
    assert Z.plan() == (3, 0)
    exact(Z, vz)
    # values set asymmetric later: back to the slice-offset layout, exact;
    # symmetric again: the stencil layout on request
    A.set_values(va)
    assert A.plan() == (3, 0)
    exact(A, va)
    A.set_values(v)
    A.set_plan(4)
    exact(A, v)
    # a structure without an entry's mirror
    keep = np.ones(rp[-1], bool)
    keep[lower[50]] = False
    rp2 = np.concatenate([[0], np.cumsum(np.add.reduceat(keep, rp[:-1]))]).astype(np.int32)
    ci2, v2 = ci[keep], v[keep]
    C = ctx.csr(rp2, ci2, v2)
    assert C.plan() == (3, 0)
    with pytest.raises(fsx.FsxError, match="structure"):
        C.set_plan(4)
    C.spmv(xd, yd)
    np.testing.assert_array_equal(yd.get(), O.Csr(rp2, ci2, v2).spmv(x))
    # alpha I + beta F on the stencil layout, bit for bit
    F = ctx.csr(rp, ci, v)
    S = F.convert(fsx.F64)
    assert S.plan() == (4, 0)
    S.shift_scale(F, 1.0, -0.37)
    vs = -0.37 * v
    vs[ci == rows] = 1.0 + vs[ci == rows]
    assert S.plan() == (4, 0)
    exact(S, vs)
    np.testing.assert_array_equal(S.download()[2], vs)


def test_slice_offsets_fallback_and_block_diagonal(ctx):
    # unsorted columns within a row: not the slice-offset layout, still exact
    rp, ci, v = P.poisson2d(20)
    ci2 = ci.copy()
    ci2[rp[5]:rp[6]] = ci2[rp[5]:rp[6]][::-1]
    v2 = v.copy()
    v2[rp[5]:rp[6]] = v2[rp[5]:rp[6]][::-1]
    A = check_spmv(ctx, rp, ci2, v2, exact=True)
    assert A.plan()[0] != 3
    with pytest.raises(fsx.FsxError):
        A.set_plan(3)
    # block-diagonal batch of stencils of different sizes: slices that
    # straddle two systems carry both offset sets
    e, rows, brp, bci, bv = P.batch(16, 4, 11, 4001, 4100)
    B = check_spmv(ctx, brp, bci, bv, exact=True)
    assert B.plan() == (3, 0)


@pytest.mark.parametrize("itype", [np.int32, np.int64])
def test_ragged_rows_both_mappings(ctx, itype):
    rp, ci, v = ragged(20_000, 1, itype=itype)
    A = check_spmv(ctx, rp, ci, v, exact=False)
    x = np.random.default_rng(9).standard_normal(A.ncols)
    ref = O.Csr(rp, ci, v).spmv(x)
    xd, yd = ctx.vec(x), ctx.zeros(A.nrows)
    # LDS-staged blocks: the long rows overflow their block's stage -> a
    # thread per row, still serial in CSR order
    A.set_plan(0)
    A.spmv(xd, yd)
    np.testing.assert_array_equal(yd.get(), ref)
    # sliced rows (heavily padded here, but exact: rows are summed serially)
    A.set_plan(2)
    A.spmv(xd, yd)
    np.testing.assert_array_equal(yd.get(), ref)
    # random columns: too many distinct offsets per slice for plan 3
    with pytest.raises(fsx.FsxError):
        A.set_plan(3)
    for lanes in (2, 4, 8, 16, 32, 64):
        A.set_plan(1, lanes)
        A.spmv(xd, yd)
        assert rel(yd.get(), ref) <= 1e-13, lanes


def test_long_rows_pick_vector_mapping(ctx):
    S = sp.random(3000, 5000, density=0.02, random_state=3, format="csr")
    A = check_spmv(ctx, S.indptr.astype(np.int32), S.indices.astype(np.int32),
                   S.data, exact=False, ncols=5000)
    kind, lanes = A.plan()
    assert (kind, lanes) == (1, 64)  # 100 nonzeros per row


def test_empty_and_rectangular(ctx):
    rp = np.zeros(11, np.int32)
    A = ctx.csr(rp, np.zeros(0, np.int32), np.zeros(0), ncols=4)
    y = ctx.vec(np.ones(10))
    A.spmv(ctx.vec(np.ones(4)), y)
    np.testing.assert_array_equal(y.get(), np.zeros(10))


def test_fp32_values(ctx):
    k = P.lognormal_field(24, 2.5, 8.0, 3001)
    rp, ci, v = P.diffusion3d(24, k)
    A64 = ctx.csr(rp, ci, v)
    A32 = A64.convert(fsx.F32)
    x = np.random.default_rng(1).standard_normal(24 ** 3)
    y = ctx.zeros(24 ** 3)
    A32.spmv(ctx.vec(x), y)   # FP32 values, FP64 x and accumulation
    ref = O.Csr(rp, ci, v.astype(np.float32).astype(np.float64)).spmv(x)
    np.testing.assert_array_equal(y.get(), ref)
    x32 = x.astype(np.float32)
    y32 = ctx.zeros(24 ** 3, fsx.F32)
    A32.spmv(ctx.vec(x32), y32)
    ref32 = sp.csr_matrix((v.astype(np.float32), ci, rp)) @ x32.astype(np.float64)
    assert rel(y32.get().astype(np.float64), ref32) < 1e-5


def test_residual_and_spmv_dot(ctx):
    k = P.lognormal_field(32, 1.0, 8.0, 7)
    rp, ci, v = P.diffusion3d(32, k)
    A = ctx.csr(rp, ci, v)
    Ao = O.Csr(rp, ci, v)
    n = 32 ** 3
    rng = np.random.default_rng(5)
    b, x = rng.standard_normal(n), rng.standard_normal(n)
    r = ctx.zeros(n)
    A.residual(ctx.vec(b), ctx.vec(x), r)
    np.testing.assert_array_equal(r.get(), Ao.residual(b, x))
    w = ctx.zeros(n)
    pw = fsx.spmv_dot(A, ctx.vec(x), w)
    wref = Ao.spmv(x)
    np.testing.assert_array_equal(w.get(), wref)
    assert abs(pw - O.dot(wref, x)) <= 1e-12 * np.abs(wref * x).sum()


def test_diag_inv_and_jacobi(ctx):
    k = P.lognormal_field(16, 1.0, 4.0, 8)
    rp, ci, v = P.diffusion3d(16, k)
    A = ctx.csr(rp, ci, v)
    Ao = O.Csr(rp, ci, v)
    d = A.diag_inv()
    Do = Ao.dinv()
    np.testing.assert_array_equal(d.get(), Do.vals)
    n = 16 ** 3
    r = np.random.default_rng(6).standard_normal(n)
    z = ctx.zeros(n)
    rd = ctx.vec(r)  # keep the handle alive across the call
    fsx.check(fsx.lib().fsx_jacobi_apply(ctx.h, d.h, rd.h, z.h))
    np.testing.assert_array_equal(z.get(), Do.spmv(r))
    # weighted Jacobi relaxation (mg/jacobi.hh:58-93), omega = float(2/3)
    b = np.random.default_rng(7).standard_normal(n)
    x0 = np.random.default_rng(8).standard_normal(n)
    xd = ctx.vec(x0)
    bd = ctx.vec(b)
    fsx.check(fsx.lib().fsx_jacobi_relax(ctx.h, A.h, np.float32(2 / 3.), 3,
                                         bd.h, xd.h))
    np.testing.assert_array_equal(xd.get(), Ao.jacobi_relax(b, x0, np.float32(2 / 3.), 3))


@pytest.mark.parametrize("plan", [3, 2, 0])
@pytest.mark.parametrize("nrelax", [1, 2, 3])
def test_jacobi_relax_layouts_bit_exact(ctx, plan, nrelax):
    """Weighted Jacobi on the slice-offset (3) and sliced (2) layouts (the
    off-diagonal row sum in CSR order, sweeps alternating between x and a
    context buffer) and on the CSR kernel (0): each equals the oracle's
    sequential sweeps bit for bit, for odd and even sweep counts."""
    n = 14
    rp, ci, v = P.diffusion3d(n, P.lognormal_field(n, 1.0, 3.0, 61))
    A = ctx.csr(rp, ci, v)
    A.set_plan(plan, 0)
    assert A.plan()[0] == plan
    Ao = O.Csr(rp, ci, v)
    b = P.uniform(n ** 3, 62)
    x0 = P.uniform(n ** 3, 63)
    xd = ctx.vec(x0)
    bd = ctx.vec(b)  # keep the handle alive across the call
    fsx.check(fsx.lib().fsx_jacobi_relax(ctx.h, A.h, np.float32(0.6666666), nrelax,
                                         bd.h, xd.h))
    np.testing.assert_array_equal(xd.get(),
                                  Ao.jacobi_relax(b, x0, np.float32(0.6666666), nrelax))


def test_transpose_product(ctx):
    rp, ci, v = ragged(3000, 4)
    A = ctx.csr(rp, ci, v)
    S = sp.csr_matrix((v, ci, rp), shape=(3000, 3000))
    x = np.random.default_rng(2).standard_normal(3000)
    y = ctx.zeros(3000)
    A.spmv_t(ctx.vec(x), y)
    assert rel(y.get(), S.T @ x) <= 1e-13


def test_rejects_bad_input(ctx):
    L = fsx.lib()
    rp = np.array([0, 2, 1], np.int32)  # not monotone
    with pytest.raises(fsx.FsxError):
        ctx.csr(rp, np.array([0, 1], np.int32), np.ones(2))
    with pytest.raises(fsx.FsxError):
        ctx.csr(np.array([0, 1], np.int32), np.array([5], np.int32), np.ones(1))
    A = ctx.csr(*P.poisson2d(4))
    with pytest.raises(fsx.FsxError):
        A.spmv(ctx.zeros(3), ctx.zeros(16))
    x = ctx.zeros(16)
    with pytest.raises(fsx.FsxError):
        A.spmv(x, x)


@pytest.mark.parametrize("name", ["sym5.mtx", "gen4.mtx"])
def test_csr_from_matrix_market(ctx, name):
    gold = Path(__file__).resolve().parent / "golden" / name
    A = fsx.csr_from_mm(ctx, gold)
    nr, nc, sym, I, J, V = O.mm_read(gold)
    rp, ci, vv = O.coo_tocsr(nr, I, J, V)
    r2, c2, v2 = A.download()
    np.testing.assert_array_equal(r2, rp)
    np.testing.assert_array_equal(c2, ci)
    np.testing.assert_array_equal(v2, vv)
    x = np.random.default_rng(1).standard_normal(nc)
    y = ctx.zeros(nr)
    A.spmv(ctx.vec(x), y)
    np.testing.assert_array_equal(y.get(), O.Csr(rp, ci, vv, ncols=nc).spmv(x))


def test_set_values_refreshes_layouts(ctx):
    n = 20
    k1 = P.lognormal_field(n, 1.0, 3.0, 1)
    k2 = P.lognormal_field(n, 1.0, 3.0, 2)
    rp, ci, v1 = P.diffusion3d(n, k1)
    _, _, v2 = P.diffusion3d(n, k2)
    A = ctx.csr(rp, ci, v1)
    A.set_plan(2)   # build the sliced layout too
    A.set_plan(3)
    A32 = A.convert(fsx.F32)
    A.set_values(v2)
    A32.set_values_from(A)
    x = np.random.default_rng(3).standard_normal(n ** 3)
    ref = O.Csr(rp, ci, v2).spmv(x)
    xd, yd = ctx.vec(x), ctx.zeros(n ** 3)
    for plan in (3, 2, 0):
        A.set_plan(plan)
        A.spmv(xd, yd)
        np.testing.assert_array_equal(yd.get(), ref, err_msg=str(plan))
    F = ctx.csr(rp, ci, v2).convert(fsx.F32)
    x32 = ctx.vec(x.astype(np.float32))
    y1, y2 = ctx.zeros(n ** 3, fsx.F32), ctx.zeros(n ** 3, fsx.F32)
    A32.spmv(x32, y1)
    F.spmv(x32, y2)
    np.testing.assert_array_equal(y1.get(), y2.get())
    with pytest.raises(fsx.FsxError):   # unrelated structure
        B32 = ctx.csr(rp, ci, v1).convert(fsx.F32)
        B32.set_values_from(A)
