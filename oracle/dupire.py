"""Oracle: implied-volatility splines and Dupire local volatility.

  * natural cubic spline, tff/math/interpolation/cubic/cubic_interpolation.py
    :163-253 (evaluation, constant outside the knots) and :256-359, :434-518
    (second derivatives from the tridiagonal system), here solved densely
    and differentiated analytically: S' = dy/h + h g'(t), S'' = linear
    interpolation of the second derivatives;
  * Interpolation2D, tff/math/interpolation/interpolation_2d/
    interpolation_2d.py:24-158: a spline in y per x row, then linear in x
    (searchsorted side='left', flat outside);
  * Dupire local volatility from implied volatilities,
    tff/experimental/local_volatility/local_volatility_model.py:102-141
    (`_dupire_local_volatility_iv`) and its precomputed-grid variant
    :144-304 (splines in S per times_grid row, row chosen left-continuously
    from times_grid[1:], dtheta/dt tabulated on the grid and splined).

The reference takes the derivatives with forward-mode autodiff; here they
are analytic.  Where the reference would differentiate through a masked
branch (outside the knots, or exactly on the last knot where its interval
has zero length) the derivative is 0, as autodiff gives there.

Pinned by (tests/test_host_dupire.py): the reference's interpolation
goldens (cubic_interpolation_test.py:167-275, interpolation_2d_test.py
:27-238), central differences of the oracle's own spline values for its
analytic derivatives, the reference's Dupire values
(local_volatility_model_test.py:192-267: 0.22848 / 0.217881 / 0.217398 at
S = 1 and 0.3 outside the strikes, with the test's 0.05 / 0.005
tolerances; 0.1 for a flat surface), the textbook identity
sigma_loc^2 = d(T sigma_imp^2)/dT for a strike-independent surface, and
the price form of Dupire's equation (local_volatility_model.py:38-99, an
independent route through Black-Scholes prices) against the textbook
implied-volatility form.
"""
import numpy as np


def spline_coeffs(x, y):
  """Half the natural spline's second derivatives at strictly increasing
  knots x (the reference's stored coefficients)."""
  x = np.asarray(x, np.float64)
  y = np.asarray(y, np.float64)
  n = x.shape[0]
  h = np.diff(x)
  a = np.zeros((n, n))
  b = np.zeros(n)
  a[0, 0] = a[-1, -1] = 1.0
  for i in range(1, n - 1):
    a[i, i - 1] = h[i - 1]
    a[i, i] = 2.0 * (h[i - 1] + h[i])
    a[i, i + 1] = h[i]
    b[i] = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1])
  m = np.linalg.solve(a, b)  # second derivatives
  return m / 2.0


def spline_eval(q, x, y, c):
  """(value, d/dq, d2/dq2) of the spline at points q."""
  q = np.asarray(q, np.float64)
  x = np.asarray(x, np.float64)
  y = np.asarray(y, np.float64)
  n = x.shape[0]
  idx = np.searchsorted(x, q, side='right') - 1
  inside = (idx >= 0) & (idx <= n - 2)
  i = np.clip(idx, 0, n - 2)
  h = x[i + 1] - x[i]
  t = (q - x[i]) / h
  c0, c1 = c[i], c[i + 1]
  dy = y[i + 1] - y[i]
  g = t * (t - 1) * ((t + 1) * 2 * c1 - (t - 2) * 2 * c0) / 6
  g1 = (2 * c1 * (3 * t * t - 1) - 2 * c0 * (3 * t * t - 6 * t + 2)) / 6
  g2 = 2 * c1 * t + 2 * c0 * (1 - t)
  val = y[i] + t * dy + h * h * g
  d1 = dy / h + h * g1
  d2 = g2
  outside_val = np.where(q > x[-1], y[-1], y[0])
  # on the last knot the reference's interval has zero length: value y[-1],
  # derivatives 0
  on_last = idx == n - 1
  val = np.where(inside, val, np.where(on_last, y[-1], outside_val))
  d1 = np.where(inside, d1, 0.0)
  d2 = np.where(inside, d2, 0.0)
  return val, d1, d2


class Surface2D:
  """Interpolation2D for one batch row: x_data [nx], y_data/z_data
  [nx, ny]."""

  def __init__(self, x_data, y_data, z_data):
    self.x = np.asarray(x_data, np.float64)
    self.y = np.asarray(y_data, np.float64)
    self.z = np.asarray(z_data, np.float64)
    self.c = [spline_coeffs(self.y[i], self.z[i])
              for i in range(self.x.shape[0])]

  def eval(self, xq, yq):
    """(value, d/dx, d/dy, d2/dy2) at paired points (xq, yq)."""
    xq, yq = np.broadcast_arrays(np.asarray(xq, np.float64),
                                 np.asarray(yq, np.float64))
    rows = [spline_eval(yq, self.y[i], self.z[i], self.c[i])
            for i in range(self.x.shape[0])]
    v = np.stack([r[0] for r in rows], -1)
    vy = np.stack([r[1] for r in rows], -1)
    vyy = np.stack([r[2] for r in rows], -1)
    nx = self.x.shape[0]
    up = np.searchsorted(self.x, xq, side='left')
    at_min, at_max = up == 0, up == nx
    lo = np.clip(up - 1, 0, nx - 1)
    hi = np.clip(up, 0, nx - 1)
    take = lambda a, k: np.take_along_axis(a, k[..., None], -1)[..., 0]
    xl, xu = self.x[lo], self.x[hi]
    w_den = np.where(at_min | at_max, xu - xl + 1, xu - xl)
    w = np.where(w_den == 0, 0.0, 1.0 / np.where(w_den == 0, 1, w_den))
    val = take(v, lo) + (xq - xl) * (take(v, hi) - take(v, lo)) * w
    dx = (take(v, hi) - take(v, lo)) * w
    dy = take(vy, lo) + (xq - xl) * (take(vy, hi) - take(vy, lo)) * w
    dyy = take(vyy, lo) + (xq - xl) * (take(vyy, hi) - take(vyy, lo)) * w
    val = np.where(at_min, take(v, lo), np.where(at_max, take(v, hi), val))
    edge = at_min | at_max
    dx = np.where(edge, 0.0, dx)
    dy = np.where(at_min, take(vy, lo), np.where(at_max, take(vy, hi), dy))
    dyy = np.where(at_min, take(vyy, lo),
                   np.where(at_max, take(vyy, hi), dyy))
    return val, dx, dy, dyy


def dupire_formula(t, s, s0, r, q, th, th_t, th_s, th_ss, no_nan=True,
                   textbook=False):
  """sigma_loc from implied vol theta and its derivatives (Dupire in
  implied-volatility form, local_volatility_model.py:121-141).

  As written in the reference the two d1 terms of the denominator carry a
  factor T; the textbook form (Gatheral 2006, eq. 1.10 in strike terms)
  has sqrt(T) there.  `textbook=False` restates the reference;
  `textbook=True` gives the textbook form (they agree at T = 1).
  """
  sq = np.sqrt(t)
  lm = np.log(s0 / s)
  with np.errstate(divide='ignore', invalid='ignore'):
    d1 = np.where(th * sq == 0, 0.0,
                  ((lm + (r - q + 0.5 * th * th) * t) / th) / sq)
  f = sq if textbook else t
  num = th * th + 2 * t * th * th_t + 2 * (r - q) * s * t * th * th_s
  den = ((1 + s * d1 * f * th_s)**2 +
         s * s * t * th * (th_ss - d1 * f * th_s * th_s))
  with np.errstate(divide='ignore', invalid='ignore'):
    var = num / den
  if no_nan:
    var = np.where(den == 0, 0.0, var)
  return np.sqrt(np.maximum(var, 0.0))


def local_vol_strike_surface(t, s, s0, r, q, surface, textbook=False):
  """Dupire local vol at (t, s) for an implied-vol surface interpolated in
  (expiry, strike) (Interpolation2D on strikes)."""
  th, th_t, th_s, th_ss = surface.eval(np.full_like(s, t), s)
  return dupire_formula(t, s, s0, r, q, th, th_t, th_s, th_ss,
                        textbook=textbook)


def local_vol_moneyness_surface(t, s, s0, r, q, surface):
  """The same for the from_market_data surface, interpolated in (expiry,
  log forward moneyness m = ln(K / S0) - (r - q) T) with constant r, q:
  theta(T, K) = I(T, m(T, K)) and the chain rule."""
  m = np.log(s / s0) - (r - q) * t
  v, i_t, i_m, i_mm = surface.eval(np.full_like(s, t), m)
  th_t = i_t - (r - q) * i_m
  th_s = i_m / s
  th_ss = (i_mm - i_m) / (s * s)
  return dupire_formula(t, s, s0, r, q, v, th_t, th_s, th_ss)


def precomputed_tables(times_grid, spot_grid, iv_fn, div_fn):
  """Rows of the precomputed variant: theta and dtheta/dt on
  times_grid x spot_grid (iv_fn(t, S) and div_fn(t, S) give them), with
  the natural spline coefficients of each row in S."""
  rows = []
  for t in times_grid:
    th = iv_fn(t, spot_grid)
    tht = div_fn(t, spot_grid)
    rows.append((th, spline_coeffs(spot_grid, th), tht,
                 spline_coeffs(spot_grid, tht)))
  return rows


def precomputed_local_vol(t, s, s0, r, q, times_grid, spot_grid, rows):
  """local_volatility_model.py:243-302: the row of times_grid[k] serves
  t in (times_grid[k], times_grid[k+1]] (left-continuous on
  times_grid[1:])."""
  k = int(np.searchsorted(np.asarray(times_grid)[1:], t, side='left'))
  th_v, th_c, tt_v, tt_c = rows[k]
  th, th_s, th_ss = spline_eval(s, spot_grid, th_v, th_c)
  th_t = spline_eval(s, spot_grid, tt_v, tt_c)[0]
  sq = np.sqrt(t)
  with np.errstate(divide='ignore', invalid='ignore'):
    ratio = np.where(s == 0, 0.0, s0 / np.where(s == 0, 1, s))
    d1 = np.where(th * sq == 0, 0.0,
                  ((np.log(ratio) + (r - q + 0.5 * th * th) * t) / th) / sq)
  num = th * th + 2 * t * th * th_t + 2 * (r - q) * s * t * th * th_s
  den = ((1 + s * d1 * t * th_s)**2 +
         s * s * t * th * (th_ss - d1 * t * th_s * th_s))
  return np.sqrt(np.maximum(num / den, 0.0))
