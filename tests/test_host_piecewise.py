"""tqf.math.piecewise against the reference's test vectors
(tests/golden/piecewise.json, transcribed from
tff/math/piecewise_test.py:31-280), plus an integral checked against
quadrature of the evaluated function."""
import numpy as np
import pytest

from conftest import load_golden
from tqf.math import piecewise


def test_find_interval_index():
  for c in load_golden('piecewise.json')['find_interval_index']:
    got = piecewise.find_interval_index(c['query'], c['lows'],
                                        last_interval_is_closed=c['closed'])
    assert got.dtype == np.int32
    np.testing.assert_array_equal(got, c['expected'])


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_values_and_integrals(dtype):
  g = load_golden('piecewise.json')
  for key in ('value_no_batch', 'value_with_batch', 'value_with_repetitions'):
    c = g[key]
    f = piecewise.PiecewiseConstantFunc(c['jumps'], c['values'], dtype=dtype)
    v = f(c['x'], left_continuous=c['left'])
    assert v.dtype == dtype
    np.testing.assert_array_equal(v, c['expected'])
  for key in ('integral_no_batch', 'integral_with_batch'):
    c = g[key]
    f = piecewise.PiecewiseConstantFunc(c['jumps'], c['values'], dtype=dtype)
    x = np.asarray(c['x'])
    np.testing.assert_allclose(f.integrate(x, x + c['width']), c['expected'],
                               rtol=1e-5, atol=1e-5)
  for key in ('matrix_event_no_batch', 'event_with_batch'):
    c = g[key]
    f = piecewise.PiecewiseConstantFunc(c['jumps'], c['values'], dtype=dtype)
    x = np.asarray(c['x'])
    np.testing.assert_allclose(f(x), c['expected_value'], rtol=1e-5,
                               atol=1e-5)
    np.testing.assert_allclose(f.integrate(x, x + c['width']),
                               c['expected_integral'], rtol=1e-5, atol=1e-5)


def test_invalid_shapes_raise():
  for c in load_golden('piecewise.json')['invalid']:
    with pytest.raises(ValueError):
      piecewise.PiecewiseConstantFunc(c['jumps'], c['values'])


def test_integral_is_the_quadrature_of_the_values():
  # midpoint quadrature of the evaluated function on a fine grid
  rs = np.random.RandomState(0)
  jumps = np.sort(rs.uniform(0, 5, (3, 6)), axis=-1)
  vals = rs.standard_normal((3, 7))
  f = piecewise.PiecewiseConstantFunc(jumps, vals, dtype=np.float64)
  a = rs.uniform(-1, 3, (3, 4))
  b = a + rs.uniform(0, 3, (3, 4))
  m = 200000
  u = (np.arange(m) + 0.5) / m
  pts = a[..., None] + (b - a)[..., None] * u  # [3, 4, m]
  quad = f(pts.reshape(3, -1)).reshape(3, 4, m).mean(-1) * (b - a)
  np.testing.assert_allclose(f.integrate(a, b), quad, atol=1e-4)
