"""CPU restatement of the reference's parallel CSR: the colouring of
topo::csr and the product of mat::parcsr.

TEST INFRASTRUCTURE ONLY: imported by tests/ and nowhere else.

Follows flecsolve/topo/csr.hh:482-618 (color: ghost lists as the sorted
unique off-block columns, force_unique; init_mats: each owned row's entries
in CSR order to the diag block with local column ids or to the offd block
with ncols_local + ghost rank) and matrices/parcsr.hh:60-91 (spmv_remote
tmp = offd x, spmv_local y = diag x, y.add(y, tmp)), with the row and
column blocks of FleCSI's equal_map (block c starts at c q + min(c, m); the
partition set by parcsr::read_mat, parcsr.hh:167-172).  FleCSI is not
vendored; equal_map is its published formula.

Each block product is compressed_ops::spmv (matrices/seq.hh:177-193): y = 0,
then y += v * x entry by entry in CSR order, one rounding per multiply and
per add (numpy does not fuse them).

Pins (tests/test_oracle_parcsr.py): at one colour the split is the identity
(diag = A, no offd, no ghosts, init_mats' one-colour case); closed forms of
the 2D five-point Laplacian's split over whole grid lines (ghost lists =
the neighbouring colours' edge lines, diag/offd counts); the blocks
reassemble the matrix entry for entry in CSR order on random ragged
matrices with unsorted and repeated columns; the product equals the serial
CSR product exactly on integer-valued data and on colour-block-diagonal
matrices, and differs from it where the reference's order does (a row
[1e16 offd, 1 diag, -1e16 offd] times ones gives 1, the serial order 0).
"""
from __future__ import annotations

import numpy as np


def equal_map(n: int, parts: int) -> np.ndarray:
    """Block bounds of FleCSI's equal_map(n, parts): parts + 1 offsets."""
    q, m = divmod(n, parts)
    return np.array([c * q + min(c, m) for c in range(parts + 1)], dtype=np.int64)


def split(rowptr, colind, vals, nrows: int, ncols: int, ncolours: int):
    """topo::csr::color + init_mats for every colour.

    Returns a list of dicts, one per colour, with rows, cols (local column
    count), diag (rowptr, colind local, vals), offd (rowptr, colind
    ncols_local + ghost rank, vals) and colmap (ghost global ids, ascending).
    """
    rowptr = np.asarray(rowptr, dtype=np.int64)
    colind = np.asarray(colind, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    rb, cb = equal_map(nrows, ncolours), equal_map(ncols, ncolours)
    out = []
    for c in range(ncolours):
        r0, r1, c0, c1 = rb[c], rb[c + 1], cb[c], cb[c + 1]
        e0, e1 = rowptr[r0], rowptr[r1]
        cols, vs = colind[e0:e1], vals[e0:e1]
        row_of = np.repeat(np.arange(r1 - r0), np.diff(rowptr[r0:r1 + 1]))
        own = (cols >= c0) & (cols < c1)          # cm.bin(cid) == colour
        ghosts = np.unique(cols[~own])            # sorted unique (force_unique)
        nloc = c1 - c0
        # boolean selection keeps CSR order inside each row
        dcnt = np.bincount(row_of[own], minlength=r1 - r0)
        ocnt = np.bincount(row_of[~own], minlength=r1 - r0)
        out.append(dict(
            rows=int(r1 - r0), cols=int(nloc),
            diag_rowptr=np.concatenate([[0], np.cumsum(dcnt)]).astype(np.int64),
            diag_colind=(cols[own] - c0).astype(np.int64),
            diag_vals=vs[own].copy(),
            offd_rowptr=np.concatenate([[0], np.cumsum(ocnt)]).astype(np.int64),
            offd_colind=(nloc + np.searchsorted(ghosts, cols[~own])).astype(np.int64),
            offd_vals=vs[~own].copy(),
            colmap=ghosts.astype(np.int64)))
    return out


def block_spmv(rowptr, colind, vals, x) -> np.ndarray:
    """compressed_ops::spmv: y_i = 0, then y_i += v x in CSR order.  The
    rows advance together one entry position at a time, so each row's sum
    is sequential."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    n = len(rowptr) - 1
    y = np.zeros(n)
    lens = np.diff(rowptr)
    for k in range(int(lens.max()) if n else 0):
        rows = np.nonzero(lens > k)[0]
        e = rowptr[rows] + k
        y[rows] = y[rows] + vals[e] * x[colind[e]]
    return y


def spmv(blocks, x, nrows: int) -> np.ndarray:
    """parcsr_ops::spmv over all colours: each colour's x on its cols space
    is [own block of x | ghost copy through colmap]."""
    x = np.asarray(x, dtype=np.float64)
    y = np.empty(nrows)
    r0, c0 = 0, 0
    for b in blocks:
        xl = np.concatenate([x[c0:c0 + b["cols"]], x[b["colmap"]]])
        tmp = block_spmv(b["offd_rowptr"], b["offd_colind"], b["offd_vals"], xl)
        yl = block_spmv(b["diag_rowptr"], b["diag_colind"], b["diag_vals"], xl)
        y[r0:r0 + b["rows"]] = yl + tmp
        r0 += b["rows"]
        c0 += b["cols"]
    return y
