"""Oracle: one-factor Hull-White short rate and the Bermudan swaption.

  * exact discretisation with constant mean reversion a and piecewise-
    constant volatility, tff/models/hull_white/vector_hull_white.py:641-781
    (step x' = e^{-a dt} x + m + sqrt(v) z, r = x + f(0, t)) with the
    integrals of :816-955: y(t) = e^{-2at} int_0^t e^{2as} sigma^2 ds, the
    conditional mean m = e^{-a t'} int_t^t' e^{as} y(s) ds and variance
    v = e^{-2a t'} int_t^t' e^{2as} sigma^2 ds (Brigo-Mercurio 2007, 3.3.1);
  * zero-coupon bonds from the short rate, :783-814:
    P(t, T) = P(0,T)/P(0,t) exp(-x G - y G^2 / 2), G = (1 - e^{-a(T-t)})/a;
  * bermudan_swaption_price (LSM), tff/models/hull_white/swaption.py:314-724;
  * the European swaption by Jamshidian's decomposition (Brigo-Mercurio
    2007, 3.11: swaption = sum of zero-bond options at the critical rate),
    used to check a one-date Bermudan, as the reference's test does
    (bermudan_swaption_test.py:176-229).

Pinned by (tests/test_host_hull_white.py): the reference's bond-price
golden (hull_white_test.py:465-485, atol 1e-12); the closed-form mean and
variance of r(t) for constant parameters (hull_white_test.py:40-57) and the
piecewise-constant-volatility standard deviation (:58-72); the integrals
m and v against numerical quadrature of their definitions; the martingale
property E[D(t) P(t, T)] = P(0, T) of simulated bonds.
"""
import numpy as np
from scipy import optimize
from scipy import special


class Params:
  """Constant mean reversion a, volatility sigma(t) piecewise constant:
  values[k] on (jumps[k-1], jumps[k]] (left-continuous)."""

  def __init__(self, a, vol_jumps, vol_values):
    self.a = float(a)
    self.jumps = np.asarray(vol_jumps, np.float64).reshape(-1)
    self.vals = np.asarray(vol_values, np.float64).reshape(-1)

  def vol(self, t):
    return self.vals[np.searchsorted(self.jumps, t, side='left')]


def _cum(p, t, f_int):
  """int_0^t of a knot-wise integrand: sum over full knot intervals plus the
  partial one; f_int(t0, t1, sigma) integrates one interval."""
  t = np.asarray(t, np.float64)
  knots = np.concatenate([[0.0], p.jumps])
  out = np.zeros_like(t)
  for i, tt in np.ndenumerate(t):
    acc = 0.0
    for k in range(knots.shape[0]):
      lo = knots[k]
      hi = knots[k + 1] if k + 1 < knots.shape[0] else np.inf
      if tt <= lo and k > 0:
        break
      acc += f_int(lo, min(tt, hi), p.vals[k] if k < p.vals.shape[0] else
                   p.vals[-1])
    out[i] = acc
  return out


def big_y(p, t):
  """int_0^t e^{2as} sigma(s)^2 ds."""
  a = p.a
  return _cum(p, t, lambda t0, t1, s: s * s / (2 * a) *
              (np.exp(2 * a * t1) - np.exp(2 * a * t0)))


def y_t(p, t):
  return np.exp(-2 * p.a * np.asarray(t)) * big_y(p, t)


def big_e(p, t):
  """int_0^t e^{as} y(s) ds, y(s) = e^{-2as} Y(s)."""
  a = p.a

  def part(t0, t1, s):
    y0 = big_y(p, np.array([t0]))[0]
    v = (np.exp(a * t1) - np.exp(a * t0) + np.exp(2 * a * t0) *
         (np.exp(-a * t1) - np.exp(-a * t0))) * s * s / (2 * a * a)
    return v + y0 * (np.exp(-a * t0) - np.exp(-a * t1)) / a
  return _cum(p, t, part)


def step_coefficients(p, times):
  """(decay, mean, variance) of the exact step between consecutive
  times (duplicates give the identity step)."""
  t = np.asarray(times, np.float64)
  a = p.a
  e_t = big_e(p, t)
  v_t = big_y(p, t)
  decay = np.exp(-a * (t[1:] - t[:-1]))
  mean = (e_t[1:] - e_t[:-1]) * np.exp(-a * t[1:])
  var = (v_t[1:] - v_t[:-1]) * np.exp(-2 * a * t[1:])
  return decay, mean, var


def sample_paths(p, times, fwd0, draws):
  """Short rates at `times` (times[0] = 0 included) from draws
  [N, len(times) - 1]: r = x + f(0, t)."""
  decay, mean, var = step_coefficients(p, times)
  sd = np.sqrt(np.maximum(var, 0.0))
  n = draws.shape[0]
  x = np.zeros(n)
  out = [x + fwd0(times[0])]
  for i in range(decay.shape[0]):
    x = (decay[i] * x + mean[i]) + sd[i] * draws[:, i]
    out.append(x + fwd0(times[i + 1]))
  return np.stack(out, 1)


def bond_price(p, r, t, big_t, zero_rate, fwd0):
  """P(t, T) given r(t) (vector_hull_white.py:783-814)."""
  a = p.a
  x = r - fwd0(t)
  g = (1 - np.exp(-a * (big_t - t))) / a
  p0t = np.exp(-zero_rate(t) * t)
  p0T = np.exp(-zero_rate(big_t) * big_t)
  return p0T / p0t * np.exp(-x * g - 0.5 * y_t(p, np.asarray(t)) * g * g)


def _zbp(a, sigma, t0, big_t, strike, p0t0, p0T):
  """Put on a zero bond maturing at T, option expiry t0 (constant a, sigma;
  Brigo-Mercurio 2007, 3.41)."""
  sp = sigma * np.sqrt((1 - np.exp(-2 * a * t0)) / (2 * a)) * (
      1 - np.exp(-a * (big_t - t0))) / a
  h = np.log(p0T / (p0t0 * strike)) / sp + sp / 2
  ncdf = lambda z: 0.5 * special.erfc(-z / np.sqrt(2))
  return strike * p0t0 * ncdf(-h + sp) - p0T * ncdf(-h)


def european_payer_swaption(a, sigma, expiry, pay_times, dcf, coupon,
                            zero_rate, notional=1.0):
  """Payer swaption (fixed coupon, floating leg worth 1 - P(t0, T_n) at
  expiry) by Jamshidian's decomposition, constant a and sigma, flat-forward
  initial curve given by zero_rate(t)."""
  pay_times = np.asarray(pay_times, np.float64)
  c = np.asarray(coupon, np.float64) * np.asarray(dcf, np.float64)
  c = c.copy()
  c[-1] += 1.0
  p = Params(a, [], [sigma])
  f0 = lambda t: zero_rate(t) + t * 0.0  # flat curve: f(0, t) = R

  def coupon_bond(r):
    return np.sum(c * bond_price(p, r, expiry, pay_times, zero_rate, f0)) - 1
  r_star = optimize.brentq(coupon_bond, -1.0, 1.0, xtol=1e-15)
  strikes = bond_price(p, r_star, expiry, pay_times, zero_rate, f0)
  p0t0 = np.exp(-zero_rate(expiry) * expiry)
  p0T = np.exp(-zero_rate(pay_times) * pay_times)
  return notional * np.sum(c * _zbp(a, sigma, expiry, pay_times, strikes,
                                    p0t0, p0T))
