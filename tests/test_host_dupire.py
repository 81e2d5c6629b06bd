"""Interpolation and Dupire local volatility (host-side surface building,
CPU torch) against the reference's golden values, the oracle and closed
forms."""
import numpy as np
import pytest
import torch

from conftest import load_golden
from oracle import dupire as odup

from tqf.math.interpolation import cubic, linear
from tqf.math.interpolation import interpolation_2d as i2d
from tqf.models import local_volatility as lvm
from tqf.experimental import volatility_surface as vsl

G = load_golden('interpolation_dupire.json')


def _np(t):
  return t.detach().cpu().numpy()


def _jvp(fn, x):
  x = torch.as_tensor(x, dtype=torch.float64)
  return torch.func.jvp(fn, (x,), (torch.ones_like(x),))[1]


# ------------------------------------------------------------ interpolation

@pytest.mark.parametrize('kind', ['clamped', 'fixed_first_derivative'])
def test_cubic_clamped_and_first_derivative_goldens(kind):
  g = G['cubic_duplicate_right']
  x = np.array(g['x_data'])
  bc = (cubic.BoundaryConditionType.CLAMPED if kind == 'clamped' else
        cubic.BoundaryConditionType.FIXED_FIRST_DERIVATIVE)
  sp = cubic.build_spline(x, g['y_data'], bc,
                          g['left_factor'] * x[..., -1:],
                          g['right_factor'] * x[..., -1:])
  np.testing.assert_almost_equal(_np(cubic.interpolate(g['x_values'], sp)),
                                 g[kind])


def test_cubic_natural_goldens():
  g = G['cubic_duplicate_natural']
  sp = cubic.build_spline(g['x_data'], g['y_data'])
  np.testing.assert_almost_equal(_np(cubic.interpolate(g['x_values'], sp)),
                                 g['expected'])
  x = np.linspace(-11, 12, 24).reshape(2, 2, 6)
  for dtype in (np.float32, np.float64):
    sp = cubic.build_spline(x, 1.0 / (1.0 + x * x), dtype=dtype)
    got = cubic.interpolate(G['cubic_batch']['search_args'], sp)
    assert got.dtype == (torch.float32 if dtype == np.float32 else
                         torch.float64)
    np.testing.assert_almost_equal(_np(got), G['cubic_batch']['expected'],
                                   decimal=6 if dtype == np.float32 else 7)
  xd = np.array([np.linspace(-5.0, 5.0, 11), np.linspace(0.0, 10.0, 11)])
  sp = cubic.build_spline(xd, 1.0 / (2.0 + xd**2))
  g = G['cubic_broadcast']
  np.testing.assert_allclose(_np(cubic.interpolate(g['x_values'], sp)),
                             g['expected'], rtol=1e-6, atol=1e-6)
  np.testing.assert_allclose(
      _np(cubic.interpolate([[g['x_values']]], sp)), [g['expected']],
      rtol=1e-6, atol=1e-6)
  with pytest.raises(ValueError):
    cubic.build_spline([[1.0, 2.0, 1.5, 3.0]], [[1.0] * 4],
                       validate_args=True)


def test_cubic_matches_oracle_and_derivatives():
  rs = np.random.RandomState(3)
  x = np.cumsum(rs.uniform(0.2, 1.0, 12))
  y = rs.standard_normal(12)
  sp = cubic.build_spline(x, y)
  c = odup.spline_coeffs(x, y)
  np.testing.assert_allclose(_np(sp.spline_coeffs), c, rtol=1e-12,
                             atol=1e-12)
  q = np.concatenate([np.linspace(x[0] - 1, x[-1] + 1, 301), x])
  v, d1, d2 = odup.spline_eval(q, x, y, c)
  f = lambda z: cubic.interpolate(z, sp)
  np.testing.assert_allclose(_np(f(torch.as_tensor(q))), v, rtol=1e-12,
                             atol=1e-12)
  np.testing.assert_allclose(_np(_jvp(f, q)), d1, rtol=1e-10, atol=1e-10)
  np.testing.assert_allclose(_np(_jvp(lambda z: _jvp(f, z), q)), d2,
                             rtol=1e-9, atol=1e-9)
  # the oracle's analytic derivatives against central differences of its
  # own values, away from knots and ends
  h = 1e-5
  inner = np.linspace(x[0] + 0.01, x[-1] - 0.01, 97)
  inner = inner[np.min(np.abs(inner[:, None] - x[None]), 1) > 1e-3]
  vp = odup.spline_eval(inner + h, x, y, c)[0]
  vm = odup.spline_eval(inner - h, x, y, c)[0]
  v0, e1, e2 = odup.spline_eval(inner, x, y, c)
  np.testing.assert_allclose((vp - vm) / (2 * h), e1, rtol=1e-6, atol=1e-6)
  np.testing.assert_allclose((vp - 2 * v0 + vm) / h**2, e2, rtol=1e-3,
                             atol=1e-3)
  # the spline reproduces the knots, and a natural spline through a line
  # is the line
  np.testing.assert_allclose(_np(f(torch.as_tensor(x))), y, rtol=1e-13,
                             atol=1e-13)
  sp2 = cubic.build_spline(x, 3 * x - 1)
  np.testing.assert_allclose(_np(cubic.interpolate(inner, sp2)),
                             3 * inner - 1, rtol=1e-12)


def test_linear_docstring_golden():
  g = G['linear_docstring']
  np.testing.assert_allclose(
      _np(linear.interpolate(g['x'], g['x_data'], g['y_data'],
                             dtype=np.float64)), g['expected'], rtol=1e-8)


def test_interpolation_2d_goldens():
  g = G['interp2d']
  td = np.array(g['times_data'])
  tv = td[:, None] * np.array(g['sigma_square_data'])
  kd = np.broadcast_to(np.array(g['strike_data'], np.float64), (5, 6))
  tol = g['tolerance']
  for dtype in (np.float32, np.float64):
    it = i2d.Interpolation2D(td, kd, tv, dtype=dtype)
    got = it.interpolate(g['times'], g['strikes'])
    assert tuple(got.shape) == (3,)
    np.testing.assert_allclose(_np(got), g['expected'], rtol=tol, atol=tol)
  tdb = np.array([g['times_data'], g['times_data_2']])
  tvb = tdb[..., None] * np.array(g['sigma_square_data_batch'])
  it = i2d.Interpolation2D(tdb, np.broadcast_to(kd, (2, 5, 6)), tvb)
  got = it.interpolate([g['times']] * 2, [g['strikes']] * 2)
  np.testing.assert_allclose(_np(got), g['expected_batch'], rtol=tol,
                             atol=tol)
  it = i2d.Interpolation2D(td, kd, tv)
  t = torch.tensor(g['times'], dtype=torch.float64)
  k = torch.tensor(g['strikes'], dtype=torch.float64)
  np.testing.assert_allclose(_np(_jvp(lambda a: it.interpolate(a, k), t)),
                             g['expected_grad_times'], rtol=tol, atol=tol)
  np.testing.assert_allclose(_np(_jvp(lambda a: it.interpolate(t, a), k)),
                             g['expected_grad_strikes'], rtol=tol, atol=tol)
  # the oracle's 2-D restatement (values and derivatives) agrees
  o = odup.Surface2D(td, kd, tv)
  v, dx, dy, _ = o.eval(np.array(g['times']), np.array(g['strikes'], float))
  np.testing.assert_allclose(v, _np(it.interpolate(t, k)), rtol=1e-12)
  np.testing.assert_allclose(dx, g['expected_grad_times'], rtol=tol,
                             atol=tol)
  np.testing.assert_allclose(dy, g['expected_grad_strikes'], rtol=tol,
                             atol=tol)


def test_interpolation_2d_closed_forms():
  # interpolation_2d_test.py:112-200: rows f1(y) = y on {0,1,2} and
  # f2(y) = y^3 on {2,3,4}; natural splines of those knots have closed forms
  f1 = lambda y: 0 if y < 0 else (y if y <= 2 else 2)

  def f2(y):
    if y < 2:
      return 8
    if y <= 3:
      return 4.5 * y**3 - 27 * y**2 + 68.5 * y - 57
    if y <= 4:
      return -4.5 * y**3 + 54 * y**2 - 174.5 * y + 186
    return 64
  it = i2d.Interpolation2D([1, 2], [[0, 1, 2], [2, 3, 4]],
                           [[0, 1, 2], [8, 27, 64]])
  ys = [-1, 0, 1, 2, 2.5, 5]
  got = it.interpolate([1] * 6 + [2] * 6, ys + ys)
  want = [f1(y) for y in ys] + [f2(y) for y in ys]
  np.testing.assert_allclose(_np(got), want, rtol=1e-4, atol=1e-4)
  got = it.interpolate([0, 1.2, 3, 4, 5, 5], ys)
  want = ([f1(ys[0]), f1(ys[1]) + 0.2 * (f2(ys[1]) - f1(ys[1]))] +
          [f2(y) for y in ys[2:]])
  np.testing.assert_allclose(_np(got), want, rtol=1e-4, atol=1e-4)


# ------------------------------------------------------------------- Dupire

def _strike_surface(strikes_row, iv_rows):
  """The reference test's surface (local_volatility_model_test.py:77-88):
  Interpolation2D on (expiry time, strike), one asset."""
  d = G['dupire']
  et = np.array(d['expiry_days'], np.float64) / 365.0
  k = np.broadcast_to(np.asarray(strikes_row, np.float64),
                      (2, len(strikes_row))) if np.ndim(strikes_row) == 1 \
      else np.asarray(strikes_row, np.float64)
  iv = np.asarray(iv_rows, np.float64)
  it = i2d.Interpolation2D(et[None], k[None], iv[None])

  def _interp(t, x):
    xt = x.transpose(0, 1)
    t = torch.as_tensor(t, dtype=xt.dtype).expand(xt.shape)
    return it.interpolate(t, xt).transpose(0, 1)
  vs = vsl.VolatilitySurface([(2020, 1, 1)],
                             vsl.dates_from_year_month_day(
                                 [[2021, 2022]], [[1, 1]], [[1, 1]]),
                             k[None], iv[None], interpolator=_interp)
  return vs, odup.Surface2D(et, k, iv), et


def _df(r):
  rt = torch.tensor([r], dtype=torch.float64)
  return lambda t: torch.exp(-rt * t)


@pytest.mark.parametrize('fn', ['iv', 'prices'])
def test_dupire_reference_values(fn):
  d = G['dupire']
  iv = np.array([d['iv_row'], d['iv_row']])
  vs, _, _ = _strike_surface(d['strikes_row'], iv)
  f = (lvm._dupire_local_volatility_iv if fn == 'iv'
       else lvm._dupire_local_volatility_prices)
  spots = torch.tensor(d['spots'], dtype=torch.float64).reshape(3, 1)
  for i, t in enumerate(d['times']):
    got = _np(f(torch.tensor(t, dtype=torch.float64), spots,
                torch.tensor([1.0], dtype=torch.float64), vs.volatility,
                _df(0.0), [0.0]))[:, 0]
    np.testing.assert_allclose(got, d['true_vols'][i], atol=d['atol'],
                               rtol=d['rtol'])
  # flat surface -> flat local volatility
  flat = np.full((2, 5), d['flat_iv'])
  vs, _, _ = _strike_surface(d['flat_strikes'], flat)
  for t in d['times']:
    got = _np(f(torch.tensor(t, dtype=torch.float64), spots,
                torch.tensor([1.0], dtype=torch.float64), vs.volatility,
                _df(0.0), [0.0]))[:, 0]
    np.testing.assert_allclose(got, d['flat_iv'], rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize('r,q', [(0.0, 0.0), (0.03, 0.01)])
def test_dupire_iv_matches_oracle_and_price_form(r, q):
  d = G['dupire']
  rs = np.random.RandomState(4)
  iv = np.stack([np.array(d['iv_row']) + 0.02 * rs.standard_normal(10),
                 np.array(d['iv_row'])[::-1] + 0.01])
  vs, osurf, _ = _strike_surface(d['strikes_row'], iv)
  s = np.linspace(0.85, 1.35, 41)
  s_t = torch.tensor(s, dtype=torch.float64).reshape(-1, 1)
  for t in (0.5, 1.3, 2.5):
    tt = torch.tensor(t, dtype=torch.float64)
    args = (torch.tensor([1.0], dtype=torch.float64), vs.volatility, _df(r),
            [q])
    got_iv = _np(lvm._dupire_local_volatility_iv(tt, s_t, *args))[:, 0]
    want = odup.local_vol_strike_surface(t, s, 1.0, r, q, osurf)
    np.testing.assert_allclose(got_iv, want, rtol=1e-9, atol=1e-12)
    # the price form is Dupire's equation through Black-Scholes prices: it
    # equals the textbook implied-volatility form, which has sqrt(T) where
    # the reference's implied-volatility form writes T (the two agree at
    # T = 1 only)
    got_px = _np(lvm._dupire_local_volatility_prices(tt, s_t, *args))[:, 0]
    book = odup.local_vol_strike_surface(t, s, 1.0, r, q, osurf,
                                         textbook=True)
    np.testing.assert_allclose(got_px, book, rtol=1e-9, atol=1e-12)
  t1 = torch.tensor(1.0, dtype=torch.float64)
  np.testing.assert_allclose(
      _np(lvm._dupire_local_volatility_iv(t1, s_t, *args)),
      _np(lvm._dupire_local_volatility_prices(t1, s_t, *args)), rtol=1e-9,
      atol=1e-12)


def test_dupire_time_only_surface_closed_form():
  # strike-independent implied vol: sigma_loc^2 = d(T theta^2)/dT exactly
  d = G['dupire']
  et = np.array(d['expiry_days']) / 365.0
  iv = np.array([[0.2] * 10, [0.3] * 10])
  vs, _, _ = _strike_surface(d['strikes_row'], iv)
  s = torch.tensor([[0.95], [1.1]], dtype=torch.float64)
  for t in (1.2, 1.5, 1.9):
    th = 0.2 + 0.1 * (t - et[0]) / (et[1] - et[0])
    want = np.sqrt(th * th + 2 * t * th * 0.1 / (et[1] - et[0]))
    for f in (lvm._dupire_local_volatility_iv,
              lvm._dupire_local_volatility_prices):
      got = _np(f(torch.tensor(t, dtype=torch.float64), s,
                  torch.tensor([1.0], dtype=torch.float64), vs.volatility,
                  _df(0.02), [0.0]))[:, 0]
      np.testing.assert_allclose(got, want, rtol=1e-6)


def test_from_market_data_surface_matches_oracle():
  # log-forward-moneyness surface of from_market_data
  d = G['dupire']
  r, q = 0.03, 0.01
  rs = np.random.RandomState(6)
  k = np.array([d['strikes_row'], d['strikes_row']])
  iv = np.stack([np.array(d['iv_row']), np.array(d['iv_row']) + 0.01 *
                 rs.standard_normal(10)])
  model = lvm.LocalVolatilityModel.from_market_data(
      1, [(2020, 1, 1)],
      vsl.dates_from_year_month_day([[2021, 2022]], [[1, 1]], [[1, 1]]),
      k[None], iv[None], [1.0], _df(r), dividend_yield=[q],
      dtype=torch.float64)
  et = np.array(d['expiry_days']) / 365.0
  m_nodes = np.log(k / 1.0) - (r - q) * et[:, None]
  osurf = odup.Surface2D(et, m_nodes, iv)
  s = np.linspace(0.9, 1.3, 17)
  for t in (0.7, 1.4, 2.2):
    got = _np(model.local_volatility_fn()(
        torch.tensor(t, dtype=torch.float64),
        torch.tensor(s, dtype=torch.float64).reshape(-1, 1)))[:, 0]
    want = odup.local_vol_moneyness_surface(t, s, 1.0, r, q, osurf)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)


def test_precomputed_dupire_matches_oracle_and_direct_formula():
  d = G['dupire']
  r = 0.02
  rs = np.random.RandomState(8)
  row = np.array(d['iv_row']) + 0.01 * rs.standard_normal(10)
  iv = np.stack([row, row])  # flat in time
  vs, osurf, _ = _strike_surface(d['strikes_row'], iv)
  tg = np.sort(np.concatenate([np.arange(0.0, 3.0, 0.15),
                               np.array(d['expiry_days']) / 365.0]))
  sg = np.array(d['strikes_row'])
  model = lvm.LocalVolatilityModel.from_volatility_surface(
      1, [1.0], vs, _df(r), dividend_yield=[0.0], times_grid=tg,
      spot_grid=sg, precompute_iv=True, dtype=torch.float64)
  fn = model.local_volatility_fn()
  rows = odup.precomputed_tables(
      tg, sg, lambda t, s: osurf.eval(np.full_like(s, t), s)[0],
      lambda t, s: osurf.eval(np.full_like(s, t), s)[1])
  s = np.linspace(0.8, 1.4, 49)
  st = torch.tensor(s, dtype=torch.float64).reshape(-1, 1)
  for t in (tg[3], 0.5 * (tg[5] + tg[6]), tg[-1]):
    got = _np(fn(torch.tensor(t, dtype=torch.float64), st))[:, 0]
    want = odup.precomputed_local_vol(t, s, 1.0, r, 0.0, tg, sg, rows)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)
    # spot_grid = the strike nodes and a time-flat surface: the row spline
    # is the surface's own spline, so the precomputed function equals the
    # direct formula
    direct = odup.local_vol_strike_surface(t, s, 1.0, r, 0.0, osurf)
    np.testing.assert_allclose(want, direct, rtol=1e-9, atol=1e-12)


def test_precomputed_table_rows_at_grid_times():
  d = G['dupire']
  vs, _, _ = _strike_surface(d['strikes_row'],
                             np.array([d['iv_row'], d['iv_row']]))
  tg = np.linspace(0.0, 2.0, 9)
  model = lvm.LocalVolatilityModel.from_volatility_surface(
      1, [1.0], vs, _df(0.01), times_grid=tg, spot_grid=d['strikes_row'],
      precompute_iv=True, dtype=torch.float64, num_spot_nodes=513)
  fn = model.local_volatility_fn()
  assert abs(fn.constant_rate() - 0.01) < 1e-14
  tab = fn.table(513)
  np.testing.assert_allclose(tab.times, tg[1:])
  x = tab.log_spot_min + tab.log_spot_step * np.arange(513)
  s = np.exp(x)
  sg = np.array(d['strikes_row'])
  s[0], s[-1] = sg[0], sg[-1] * (1 - 1e-13)
  outside = np.array([[sg[0] * 0.5], [sg[-1] * 2.0]])
  for k in (0, 4, 7):
    t = torch.tensor(tg[k + 1], dtype=torch.float64)
    want = _np(fn(t, torch.tensor(s, dtype=torch.float64).reshape(-1, 1)))
    np.testing.assert_allclose(tab.vols[k], want[:, 0], rtol=1e-14)
    # the edge values are the function outside spot_grid (constant in S)
    o = _np(fn(t, torch.tensor(outside, dtype=torch.float64)))[:, 0]
    np.testing.assert_allclose(tab.outside[k], o, rtol=1e-14)
    # and the host lookup of the table reproduces the function between
    # nodes to the linear-interpolation error of the ln S axis
    q = np.exp(np.linspace(np.log(0.7), np.log(1.5), 301))
    q = q[(np.abs(q / sg[0] - 1) > 1e-6) & (np.abs(q / sg[-1] - 1) > 1e-6)]
    exact = _np(fn(t, torch.tensor(q, dtype=torch.float64).reshape(-1, 1)))
    got = _np(tab(t, torch.tensor(q, dtype=torch.float64).reshape(-1, 1)))
    np.testing.assert_allclose(got, exact, rtol=2e-5)
