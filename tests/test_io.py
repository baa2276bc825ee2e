"""Host input path (Matrix Market reader, COO -> CSR) of libfsx against the
oracle's restatement of matrix_market.hh / coo::tocsr; host-only calls, no
GPU.  Bar: bit-exact indices, values and flags."""
from pathlib import Path

import numpy as np
import pytest

from flecsolve_amd import fsx
from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", ["sym5.mtx", "gen4.mtx"])
def test_mm_read_matches_oracle(name):
    got = fsx.mm_read(GOLD / name)
    ref = O.mm_read(GOLD / name)
    assert got[:3] == ref[:3]
    for g, r in zip(got[3:], ref[3:]):
        np.testing.assert_array_equal(g, r)


def test_mm_symmetric_mirror_order_and_float_rounding():
    nr, nc, sym, I, J, V = fsx.mm_read(GOLD / "sym5.mtx")
    assert (nr, nc, sym) == (5, 5, True)
    assert len(I) == 15
    assert (I[1], J[1], I[2], J[2]) == (1, 0, 0, 1)
    assert V[0] == float(np.float32(4.1)) != 4.1


def test_tocsr_matches_oracle_random():
    rng = np.random.default_rng(5)
    for nrows, nnz in [(1, 0), (7, 40), (300, 20000), (50_000, 200_000)]:
        I = rng.integers(0, nrows, nnz)
        J = rng.integers(0, 2 * nrows, nnz)
        V = rng.standard_normal(nnz)
        got = fsx.coo_tocsr(nrows, I, J, V)
        ref = O.coo_tocsr(nrows, I, J, V)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)


def write(tmp_path, text, name="m.mtx"):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_symmetric_flag_in_comment_and_empty_rows(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real general\n"
                        "% this file is symmetric after all\n4 4 3\n1 1 2.5\n4 1 -1\n4 4 3\n")
    got, ref = fsx.mm_read(p), O.mm_read(p)
    assert got[2] is True and got[:3] == ref[:3]
    for g, r in zip(got[3:], ref[3:]):
        np.testing.assert_array_equal(g, r)
    rp, ci, vv = fsx.coo_tocsr(got[0], *got[3:])
    np.testing.assert_array_equal(rp, [0, 2, 2, 2, 4])  # rows 1, 2 empty


def test_header_only_and_errors(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real general\n3 3 0\n")
    nr, nc, sym, I, J, V = fsx.mm_read(p)
    assert (nr, nc, sym, I.size) == (3, 3, False, 0)
    with pytest.raises(fsx.FsxError):
        fsx.mm_read(tmp_path / "missing.mtx")
    with pytest.raises(fsx.FsxError):   # fewer entry lines than announced
        fsx.mm_read(write(tmp_path, "%%MatrixMarket\n2 2 3\n1 1 1\n", "short.mtx"))
    with pytest.raises(fsx.FsxError):   # two spaces: an empty token, as stoi would reject
        fsx.mm_read(write(tmp_path, "%%MatrixMarket\n2 2 1\n1  1 1\n", "bad.mtx"))
    with pytest.raises(fsx.FsxError):
        fsx.coo_tocsr(2, np.array([0, 2]), np.array([0, 0]), np.array([1.0, 1.0]))
