"""Hull-White: the oracle pinned by closed forms, quadrature and the
reference's bond-price golden; the model's host-side step constants and
bond constants against the oracle."""
import numpy as np
import pytest
import torch
from scipy import integrate

from oracle import hull_white as ohw

from tqf.math.piecewise import PiecewiseConstantFunc
from tqf.models.hull_white.one_factor import HullWhiteModel1F

FLAT = lambda t: 0.01 * torch.ones_like(torch.as_tensor(t))
R01 = lambda t: 0.01 + 0.0 * np.asarray(t, np.float64)


def test_bond_price_golden():
  # hull_white_test.py:465-485: a = 0.1, sigma = 0.01, flat 1% curve
  p = ohw.Params(0.1, [], [0.01])
  got = ohw.bond_price(p, np.array([0.011, 0.01]), np.array([1.0, 2.0]),
                       np.array([2.0, 3.5]), R01, R01)
  np.testing.assert_allclose(got, [0.98906753, 0.98495442], rtol=1e-6,
                             atol=1e-12)
  m = HullWhiteModel1F(PiecewiseConstantFunc([], [0.1]),
                       PiecewiseConstantFunc([], [0.01]), FLAT,
                       dtype=torch.float64)
  f0, p0big, g, y = m._bond_constants(np.array([1.0, 2.0]),
                                      np.array([2.0, 3.5]))
  mine = p0big * np.exp(-(np.array([0.011, 0.01]) - f0) * g - 0.5 * y * g * g)
  np.testing.assert_allclose(mine, got, rtol=1e-14)


def _quad_moments(a, jumps, vals, t0, t1):
  """Conditional mean and variance of the exact step by quadrature of their
  definitions."""
  p = ohw.Params(a, jumps, vals)
  sig2 = lambda s: float(p.vol(np.array([s]))[0])**2
  pts = [j for j in jumps if 0 < j < t1]
  big_y = lambda s: integrate.quad(lambda u: np.exp(2 * a * u) * sig2(u), 0,
                                   s, points=[j for j in pts if j < s] or None,
                                   epsabs=1e-15, epsrel=1e-13)[0]
  y = lambda s: np.exp(-2 * a * s) * big_y(s)
  inner = [j for j in pts if t0 < j < t1] or None
  mean = np.exp(-a * t1) * integrate.quad(lambda s: np.exp(a * s) * y(s), t0,
                                          t1, points=inner, epsabs=1e-15,
                                          epsrel=1e-12)[0]
  var = np.exp(-2 * a * t1) * integrate.quad(
      lambda s: np.exp(2 * a * s) * sig2(s), t0, t1, points=inner,
      epsabs=1e-15, epsrel=1e-12)[0]
  return mean, var


@pytest.mark.parametrize('a,jumps,vals', [(0.1, [], [0.01]),
                                          (0.03, [0.1, 2.0],
                                           [0.01, 0.02, 0.01])])
def test_step_constants_against_quadrature(a, jumps, vals):
  p = ohw.Params(a, jumps, vals)
  t = np.array([0.0, 0.05, 0.1, 0.7, 2.0, 2.5, 3.0])
  decay, mean, var = ohw.step_coefficients(p, t)
  for i in range(t.shape[0] - 1):
    qm, qv = _quad_moments(a, jumps, vals, t[i], t[i + 1])
    np.testing.assert_allclose(mean[i], qm, rtol=1e-8, atol=1e-16)
    np.testing.assert_allclose(var[i], qv, rtol=1e-8, atol=1e-16)
  np.testing.assert_allclose(decay, np.exp(-a * np.diff(t)), rtol=1e-15)


def _propagate(decay, mean, var):
  m = v = 0.0
  for e, mu, s2 in zip(decay, mean, var):
    m = e * m + mu
    v = e * e * v + s2
  return m, v


def test_moments_closed_form():
  # hull_white_test.py:40-57 (Brigo-Mercurio 3.3.1): with x(0) = 0,
  # E[x_t] = sigma^2 / (2 a^2) (1 - e^{-at})^2, Var = sigma^2 (1 - e^{-2at}) / 2a
  a, s = 0.1, 0.01
  p = ohw.Params(a, [], [s])
  t = np.concatenate([[0.0], np.linspace(0.1, 1.0, 7), [1.7, 3.0]])
  m, v = _propagate(*ohw.step_coefficients(p, t))
  tt = t[-1]
  np.testing.assert_allclose(m, s * s / 2 / a**2 * (1 - np.exp(-a * tt))**2,
                             rtol=1e-12)
  np.testing.assert_allclose(v, s * s / 2 / a * (1 - np.exp(-2 * a * tt)),
                             rtol=1e-12)
  # piecewise-constant volatility: the reference test's standard deviation
  # (hull_white_test.py:58-72)
  k, vols, knots = 0.1, [0.01, 0.02, 0.01], [0.0, 0.1, 2.0]
  p = ohw.Params(k, [0.1, 2.0], vols)
  for tt in (0.1, 1.0, 2.0, 3.0):
    grid = np.unique(np.concatenate([[0.0, tt], [0.1, 2.0]]))
    grid = grid[grid <= tt]
    _, v = _propagate(*ohw.step_coefficients(p, grid))
    var = 0.0
    for j in range(len(knots)):
      lo = knots[j]
      hi = knots[j + 1] if j + 1 < len(knots) else np.inf
      if tt <= lo:
        break
      var += vols[j]**2 / 2 / k * (np.exp(2 * k * min(tt, hi)) -
                                   np.exp(2 * k * lo))
    np.testing.assert_allclose(np.sqrt(v), np.exp(-k * tt) * np.sqrt(var),
                               rtol=1e-12)


def test_model_host_constants_match_oracle():
  p = ohw.Params(0.07, [0.5, 2.0], [0.01, 0.02, 0.015])
  m = HullWhiteModel1F(0.07, PiecewiseConstantFunc([0.5, 2.0],
                                                   [0.01, 0.02, 0.015]),
                       FLAT, dtype=torch.float64)
  t = np.array([0.0, 0.3, 0.5, 0.5, 1.0, 2.0, 2.5, 4.0])
  decay, mean, sd = m._step_constants(t)
  od, om, ov = ohw.step_coefficients(p, t)
  np.testing.assert_allclose(decay, od, rtol=1e-14)
  np.testing.assert_allclose(mean, om, rtol=1e-12, atol=1e-18)
  np.testing.assert_allclose(sd, np.sqrt(np.maximum(ov, 0)), rtol=1e-12)
  np.testing.assert_allclose(m._compute_yt(t), ohw.y_t(p, t), rtol=1e-12,
                             atol=1e-18)
  np.testing.assert_allclose(m._fwd0(t), 0.01, rtol=1e-14)
  # bond constants on a broadcast (times x tenors) grid
  tk = np.array([0.3, 1.0, 2.0])[None, :]
  tau = np.array([0.0, 0.5, 1.0])[:, None]
  f0, p0big, g, y = m._bond_constants(tk, tk + tau)
  assert f0.shape == p0big.shape == g.shape == y.shape == (3, 3)
  np.testing.assert_allclose(g, (1 - np.exp(-0.07 * tau)) / 0.07 + 0 * tk,
                             rtol=1e-14, atol=1e-16)
  all_t, mask = m._prepare_grid(np.array([0.3, 1.0, 2.0]), None)
  np.testing.assert_allclose(all_t, [0.0, 0.3, 0.5, 1.0, 2.0, 2.0])
  assert mask.tolist() == [False, True, False, True, True, False]


def test_jamshidian_reference_values():
  # swaption_test.py:97-123, 300-316 (1y x 1y payer, quarterly, a = 0.03,
  # sigma = 0.02, flat 1%): the reference's analytic values; it finds the
  # critical rate with its brentq default tolerance, which leaves a few
  # 1e-7 of relative error in its 1y value
  pay = np.array([1.25, 1.5, 1.75, 2.0])
  for expiry, want, rtol in ((0.5, 1.02849462, 3e-8),
                             (1.0, 0.71632434, 1e-6)):
    got = 100 * ohw.european_payer_swaption(0.03, 0.02, expiry, pay,
                                            np.full(4, 0.25),
                                            np.full(4, 0.011), R01)
    np.testing.assert_allclose(got, want, rtol=rtol)


def test_jamshidian_limits():
  # sigma -> 0: the payer swaption is worth its intrinsic value on the
  # initial curve; the zero-bond put satisfies put-call parity
  pay = np.array([1.5, 2.0, 2.5, 3.0])
  dcf = np.full(4, 0.5)
  for cpn in (0.005, 0.011, 0.02):
    v = ohw.european_payer_swaption(0.03, 1e-7, 1.0, pay, dcf,
                                    np.full(4, cpn), R01)
    p0 = np.exp(-0.01 * np.concatenate([[1.0], pay]))
    intrinsic = max(p0[0] - p0[-1] - cpn * 0.5 * p0[1:].sum(), 0.0)
    np.testing.assert_allclose(v, intrinsic, atol=1e-9)
  a, s, t0, big_t, k = 0.03, 0.01, 1.0, 3.0, 0.97
  p0t0, p0T = np.exp(-0.01 * t0), np.exp(-0.01 * big_t)
  put = ohw._zbp(a, s, t0, big_t, k, p0t0, p0T)
  sp = s * np.sqrt((1 - np.exp(-2 * a * t0)) / (2 * a)) * (
      1 - np.exp(-a * (big_t - t0))) / a
  h = np.log(p0T / (p0t0 * k)) / sp + sp / 2
  from scipy import special
  ncdf = lambda z: 0.5 * special.erfc(-z / np.sqrt(2))
  call = p0T * ncdf(h) - k * p0t0 * ncdf(h - sp)
  np.testing.assert_allclose(call - put, p0T - k * p0t0, rtol=1e-12)
