"""Oracle SpMV, COO->CSR, Matrix Market reader, Dinv and Jacobi relaxation."""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"


def rand_csr(n, m, density, seed, itype=np.int32):
    S = sp.random(n, m, density=density, random_state=seed, format="csr")
    S.data = np.random.default_rng(seed).standard_normal(S.nnz)
    return S.indptr.astype(itype), S.indices.astype(itype), S.data, S


@pytest.mark.parametrize("itype", [np.int32, np.int64])
def test_spmv_matches_scipy(itype):
    rp, ci, v, S = rand_csr(300, 300, 0.05, 3, itype)
    A = O.Csr(rp, ci, v)
    x = np.random.default_rng(4).standard_normal(300)
    np.testing.assert_allclose(A.spmv(x), S @ x, rtol=1e-13, atol=1e-13)


def test_spmv_serial_order_exact():
    # CSR-order accumulation from 0 (matrices/seq.hh:183-188): an
    # exact-integer matrix gives exact results, and a cancelling row exposes
    # the order of summation.
    rp = np.array([0, 3, 5], np.int64)
    ci = np.array([0, 1, 2, 0, 2], np.int64)
    v = np.array([1e16, 1.0, -1e16, 2.0, 3.0])
    A = O.Csr(rp, ci, v, ncols=3)
    y = A.spmv(np.array([1.0, 1.0, 1.0]))
    assert y[0] == ((0.0 + 1e16) + 1.0) - 1e16  # = 0.0 in serial order
    assert y[1] == 5.0


def test_residual_is_b_minus_ax():
    rp, ci, v, S = rand_csr(50, 50, 0.2, 5)
    A = O.Csr(rp, ci, v)
    b = np.arange(50.0)
    x = np.linspace(-1, 1, 50)
    np.testing.assert_array_equal(A.residual(b, x), b - A.spmv(x))


def test_coo_tocsr_brute_force():
    rng = np.random.default_rng(0)
    nrows = 7
    nnz = 40
    I = rng.integers(0, nrows, nnz)
    J = rng.integers(0, 9, nnz)
    V = rng.standard_normal(nnz)
    rp, ci, vv = O.coo_tocsr(nrows, I, J, V)
    # stable sort by row only; input order kept inside a row; duplicates kept
    order = sorted(range(nnz), key=lambda e: I[e])  # Python sort is stable
    np.testing.assert_array_equal(ci, J[order])
    np.testing.assert_array_equal(vv, V[order])
    np.testing.assert_array_equal(rp, np.concatenate([[0], np.cumsum(np.bincount(I, minlength=nrows))]))


def test_mm_read_symmetric_mirrors_and_rounds_through_float():
    nr, nc, sym, I, J, V = O.mm_read(GOLD / "sym5.mtx")
    assert (nr, nc, sym) == (5, 5, True)
    # each off-diagonal entry is followed by its mirror (matrix_market.hh:95-99)
    assert len(I) == 5 + 2 * 5
    assert (I[1], J[1]) == (1, 0) and (I[2], J[2]) == (0, 1)
    assert V[1] == V[2] == float(np.float32(-1.3))
    assert V[0] == float(np.float32(4.1)) and V[0] != 4.1


def test_mm_read_general_and_tocsr():
    nr, nc, sym, I, J, V = O.mm_read(GOLD / "gen4.mtx")
    assert not sym and len(I) == 7
    rp, ci, vv = O.coo_tocsr(nr, I, J, V)
    np.testing.assert_array_equal(rp, [0, 3, 4, 6, 7])
    # row 0 keeps file order: (1,1), (1,3), (1,2)
    np.testing.assert_array_equal(ci[:3], [0, 2, 1])
    assert vv[1] == float(np.float32(1 / 3))


def test_dinv_csr():
    rp, ci, v, S = rand_csr(40, 40, 0.2, 7)
    S = (S + sp.eye(40) * 5).tocsr()
    A = O.Csr(S.indptr.astype(np.int64), S.indices.astype(np.int64), S.data)
    D = A.dinv()
    r = np.random.default_rng(2).standard_normal(40)
    np.testing.assert_array_equal(D.spmv(r), 0.0 + (1.0 / S.diagonal()) * r)


def test_jacobi_relax_by_hand_and_fixed_point():
    # 3x3 by hand: one sweep from x = 0 gives omega * b / diag
    S = sp.csr_matrix(np.array([[4.0, -1, 0], [-1, 4, -1], [0, -1, 4]]))
    A = O.Csr(S.indptr.astype(np.int64), S.indices.astype(np.int64), S.data)
    b = np.array([1.0, 2.0, 3.0])
    om = np.float32(2 / 3.)
    x1 = A.jacobi_relax(b, np.zeros(3), om, 1)
    np.testing.assert_allclose(x1, float(om) * b / 4.0, rtol=1e-15)
    # the exact solution is a fixed point of the relaxation
    xs = np.linalg.solve(S.toarray(), b)
    np.testing.assert_allclose(A.jacobi_relax(b, xs, om, 3), xs, rtol=1e-14)
    # many sweeps converge to it
    np.testing.assert_allclose(A.jacobi_relax(b, np.zeros(3), om, 200), xs, rtol=1e-12)
