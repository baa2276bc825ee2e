"""Oracle BLAS-1 and reductions against the reference's closed forms."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import vector_kat


class OracleOps:
    """Adapter: handles are mutable numpy arrays, ops call the oracle."""

    def new(self, v):
        return np.array(v, dtype=np.float64)

    def get(self, h):
        return h

    def _set(self, dst, val):
        dst[:] = val

    def add(self, z, x, y): self._set(z, O.binop("add", x, y))
    def subtract(self, z, x, y): self._set(z, O.binop("subtract", x, y))
    def multiply(self, z, x, y): self._set(z, O.binop("multiply", x, y))
    def divide(self, z, x, y): self._set(z, O.binop("divide", x, y))
    def add_scalar(self, y, x, a): self._set(y, O.unop("add_scalar", x, a))
    def scale_to(self, y, a, x): self._set(y, O.unop("scale_to", x, a))
    def reciprocal(self, y, x): self._set(y, O.unop("reciprocal", x))
    def linear_sum(self, z, a, x, b, y): self._set(z, O.linear_sum(a, x, b, y))
    def axpy(self, z, a, x, y): self._set(z, O.axpy(a, x, y))
    def copy(self, d, s): self._set(d, s)
    def axpby(self, z, a, b, x): self._set(z, O.axpby(a, b, x, z))
    def abs(self, y, x): self._set(y, O.unop("abs", x))
    def set_scalar(self, x, a): x[:] = a
    def min(self, x): return O.vmin(x)
    def max(self, x): return O.vmax(x)
    def dot(self, x, y): return O.dot(x, y)
    def l1norm(self, x): return O.l1norm(x)
    def l2norm(self, x): return O.l2norm(x)


def test_flecsi_vector_closed_forms():
    vector_kat.run(OracleOps())


def test_serial_sum_order():
    # scalar_prod accumulates left to right (vectors/seq.hh:282-295): a
    # cancellation-sensitive input distinguishes it from pairwise summation.
    x = np.array([1e16, 1.0, -1e16, 1.0])
    assert O.dot(x, np.ones(4)) == ((((0.0 + 1e16) + 1.0) - 1e16) + 1.0)
    assert O.l2norm(np.array([3.0, 4.0])) == 5.0


def test_inf_norm_quirk():
    # inf_norm starts from numeric_limits<double>::min() (vectors/seq.hh:251)
    assert O.inf_norm(np.zeros(4)) == np.finfo(np.float64).tiny
    assert O.inf_norm(np.array([-3.0, 2.0])) == 3.0
    assert O.vmin(np.array([])) == np.finfo(np.float64).max
    assert O.vmax(np.array([])) == -np.finfo(np.float64).max


def test_set_random_is_mt19937_canonical():
    # libstdc++ uniform_real_distribution<double>(0,1) over mt19937 consumes
    # two 32-bit draws per value: (a + b * 2^32) / 2^64.  numpy's MT19937 with
    # the legacy seeding `RandomState(seed)` uses the same init_genrand.
    rs = np.random.RandomState(7)
    raw = rs.randint(0, 2**32, size=8, dtype=np.uint64)
    ref = (raw[0::2] + raw[1::2] * 2.0**32) / 2.0**64
    got = O.set_random(4, 7)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("alpha", [0.0, -2.5, 3.0])
def test_axpy_rounding_matches_expression(alpha):
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(1000), rng.standard_normal(1000)
    # no fused multiply-add: alpha*x rounded, then + y rounded
    np.testing.assert_array_equal(O.axpy(alpha, x, y), alpha * x + y)


def test_lp_powsum_f32_folds_in_float():
    """FP32 lp_norm_local (topo_tasks.hh:216-226) folds std::pow(float, int)
    terms -- evaluated in double -- into a float.  x = 1 + 2^-10: x^3 =
    1 + 3 2^-10 + 3 2^-20 + 2^-30 in double, 1 + 3 2^-10 + 3 2^-20 once
    rounded to float; 16 of them sum exactly in float (21 significant bits),
    so the fold is 16 (1 + 3 2^-10 + 3 2^-20), not the double sum's
    16 (1 + 3 2^-10 + 3 2^-20 + 2^-30).  Random input: the fold is the
    serial float sum of float-rounded terms, in index order."""
    x = np.full(16, 1 + 2.0 ** -10, np.float32)
    t = 1 + 3 * 2.0 ** -10 + 3 * 2.0 ** -20
    assert O.lp_powsum_f32(x, 3) == 16 * t
    assert 16 * t != 16 * (1 + 2.0 ** -10) ** 3
    r = np.random.default_rng(5).random(1000).astype(np.float32) + np.float32(0.5)
    terms = np.power(r.astype(np.float64), 3).astype(np.float32)
    assert O.lp_powsum_f32(r, 3) == float(np.cumsum(terms, dtype=np.float32)[-1])
