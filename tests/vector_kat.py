"""Closed-form checks of every BLAS-1 operation, shared by the oracle and the
device tests.

This is the sequence of flecsolve/vectors/test/flecsi_vector.cc:309-402 on a
32-entry mesh whose fields hold multiples of the global id
(x = gid, y = 2 gid, z = 3 gid; flecsi_vector.cc:21-24, 41-58), with the
answers of its *_check structs (flecsi_vector.cc:99-307) and its tolerance
(sum |answer - got| < n * 1e-8).  `ops` adapts the storage under test: it
exposes new(values) -> handle, get(handle) -> ndarray and the vec::core
method names taking handles.
"""
import numpy as np

N = 32
FTOL = 1e-8


def run(ops):
    gid = np.arange(N, dtype=np.float64)
    x = ops.new(1 * gid)
    y = ops.new(2 * gid)
    z = ops.new(3 * gid)
    tmp = ops.new(np.zeros(N))
    results = {}

    def check(name, handle, answer):
        err = np.abs(answer - ops.get(handle)).sum()
        results[name] = err
        assert err < N * FTOL, (name, err)

    ops.add(tmp, x, z)
    check("add", tmp, gid + 3 * gid)
    ops.subtract(tmp, x, z)
    check("subtract", tmp, gid - 3 * gid)
    ops.multiply(tmp, x, z)
    check("multiply", tmp, gid * 3 * gid)
    ops.add_scalar(x, x, 1.0)
    check("add scalar", x, gid + 1)
    ops.divide(tmp, y, x)
    check("divide", tmp, 2 * gid / (gid + 1))
    ops.add_scalar(x, x, -1.0)
    ops.scale_to(tmp, 2.0, x)
    check("scale", tmp, gid * 2)
    ops.add_scalar(y, y, 1.0)
    ops.reciprocal(tmp, y)
    check("reciprocal", tmp, 1.0 / (2 * gid + 1))
    ops.add_scalar(y, y, -1.0)
    ops.linear_sum(tmp, 8.0, y, 9.0, z)
    check("linear sum", tmp, 2 * gid * 8 + 3 * gid * 9)
    ops.axpy(tmp, 7.0, x, y)
    check("axpy", tmp, gid * 7 + 2 * gid)
    ops.copy(tmp, y)
    ops.axpby(tmp, 4.0, 11.0, z)
    check("axpby", tmp, 3 * gid * 4 + 2 * gid * 11)
    ops.add_scalar(tmp, y, -4.0)
    ops.abs(tmp, tmp)
    check("abs", tmp, np.abs(2 * gid - 4))
    ops.add_scalar(tmp, y, -7.0)
    assert ops.min(tmp) == -7
    assert ops.max(z) == 93
    # reductions (flecsi_vector.cc:382-392)
    ops.set_scalar(x, 1.5)
    ops.set_scalar(y, 3.8)
    assert abs(32 * 1.5 * 3.8 - ops.dot(x, y)) < FTOL
    ops.set_scalar(x, -3.141719)
    assert abs(32 * 3.141719 - ops.l1norm(x)) < FTOL
    assert abs(np.sqrt(32 * 3.141719 * 3.141719) - ops.l2norm(x)) < FTOL
    return results
