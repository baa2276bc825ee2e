"""CPU model of fsx_ir_cg_rr_solve, the FP32 CG with FP64 residual
replacement (no reference counterpart), used as the checker of its device
implementation.

It follows the device's operation order and precisions: the FP32 matrix
S A S (a_ij (s_i s_j) rounded once, s = 1/sqrt(a_ii)); FP32 vectors whose
updates are formed in FP64 and rounded once; FP64 scalars and dot products;
the single-reduction recurrences (alpha = rho / p.w, the next rho estimated
as alpha^2 w.w - rho for beta, the recursive |r|^2 summed by the update for
the next alpha and the tolerance test); segments that end at
max(seg_rtol |rh_0|, 0.5 rtol |b| |rh_0| / |r|) or after maxiter updates;
between segments x += c S y (FP64), r = b - A x (FP64), the replacement
rh = (float)(s (r / c)), p += rh_new - rh_old, y = 0.

`keep_p=False` restarts the FP32 CG at every replacement (p = rh): the
restarted scheme the replacement is measured against.
`single_reduction=False` takes beta from the updated |r|^2 (the batched
solver's two-reduction form) instead of the estimate; `acc64=True` sums each
row of the FP32 product in FP64 before rounding it (the batched solver's
product; the global FP32 product sums in FP32).  `low=np.float64`
runs the same recurrences with FP64 vectors and matrix: then the iteration
is CG with Jacobi preconditioning in FP64 (S A S y = S b, x = S y), which
the tests hold against the oracle's CG + Dinv.
"""
import numpy as np
import scipy.sparse as sp

F32 = np.float32


def scaled_matrix(rp, ci, v, low=F32):
    """(s, S A S as a CSR matrix of dtype `low`) for a CSR matrix with a
    positive stored diagonal."""
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    d = np.zeros(n)
    on = ci == rows
    d[rows[on]] = v[on]
    s = 1.0 / np.sqrt(d)
    vals = (v * (s[rows] * s[ci])).astype(low)
    return s, sp.csr_matrix((vals, ci, rp), shape=(n, n))


def _dot(a, b):
    return float(np.dot(a.astype(np.float64), b.astype(np.float64)))


def rr_solve(rp, ci, v, b, rtol, seg_rtol, maxiter, max_segments, x0=None,
             keep_p=True, low=F32, single_reduction=True, acc64=False):
    """Returns (x, segments, FP32 iterations, status, true |r|/|b|); status
    1 converged, 4 max_segments reached, 5 the FP32 residual vanished."""
    n = rp.size - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    s, Ah = scaled_matrix(rp, ci, v, low)
    Ah64 = Ah.astype(np.float64) if acc64 else None
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    bn = np.linalg.norm(b)
    bn = bn if bn > 0 else 1.0
    r = b - A @ x
    rnorm = np.linalg.norm(r)
    if rnorm / bn <= rtol or rnorm == 0:
        return x, 0, 0, 1, rnorm / bn
    c = rnorm
    rh = (s * (r * (1.0 / c))).astype(low)
    p = rh.copy()
    y = np.zeros(n, low)
    rho = _dot(rh, rh)
    segs = iters = 0
    while True:
        if segs >= max_segments:
            return x, segs, iters, 4, rnorm / bn
        if not rho > 0:
            return x, segs, iters, 5, rnorm / bn
        rh0 = np.sqrt(rho)
        tol = max(seg_rtol * rh0, 0.5 * rtol * bn * rh0 / rnorm)
        it = 0
        while True:
            if it > 0 and (np.sqrt(rho) < tol or it >= maxiter):
                break
            if acc64:
                w = (Ah64 @ p.astype(np.float64)).astype(low)
            else:
                w = (Ah @ p).astype(low)
            pw, ww = _dot(w, p), _dot(w, w)
            alpha = rho / pw
            p64 = p.astype(np.float64)
            rh = ((-alpha) * w.astype(np.float64) + rh.astype(np.float64)).astype(low)
            rho_new = _dot(rh, rh)
            # single reduction: beta from the estimate alpha^2 w.w - rho, known
            # before the update; two reductions: from the updated |r|^2
            beta = ((alpha * alpha * ww - rho) if single_reduction else rho_new) / rho
            y = (alpha * p64 + y.astype(np.float64)).astype(low)
            p = (beta * p64 + rh.astype(np.float64)).astype(low)
            rho = rho_new
            it += 1
        iters += it
        segs += 1
        x = x + c * (s * y.astype(np.float64))
        r = b - A @ x
        rnorm = np.linalg.norm(r)
        if rnorm / bn <= rtol or rnorm == 0:
            return x, segs, iters, 1, rnorm / bn
        new = (s * (r * (1.0 / c))).astype(low)
        if keep_p:
            p = (p.astype(np.float64) + (new.astype(np.float64) - rh.astype(np.float64))).astype(low)
        else:
            p = new.copy()
        rh = new
        y = np.zeros(n, low)
        rho = _dot(rh, rh)
