"""Hull-White on the GPU: the fused exact-step kernel and bond
reconstitution against the oracle on the same draws, moments against
closed forms, and Bermudan swaptions against the reference's values
(bermudan_swaption_test.py:85-293) and the oracle's pipeline."""
import numpy as np
import pytest
import torch

from oracle import hull_white as ohw
from oracle import lsm as olsm

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
  import tqf
  from tqf.math.piecewise import PiecewiseConstantFunc
  from tqf.math.random.multivariate_normal import RandomType
  from tqf.models.hull_white import swaption as sw

R01 = lambda t: 0.01 + 0.0 * np.asarray(t, np.float64)


def _flat(t):
  return 0.01 * torch.ones_like(torch.as_tensor(t))


def _np(t):
  return t.detach().cpu().numpy()


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12),
                                       (torch.float32, 2e-5)])
def test_paths_and_bonds_match_oracle(dtype, tol):
  npd = np.float64 if dtype == torch.float64 else np.float32
  vol = PiecewiseConstantFunc([0.5, 2.0], [0.01, 0.02, 0.015])
  model = tqf.models.hull_white.HullWhiteModel1F(0.07, vol, _flat,
                                                 dtype=dtype)
  times = np.array([0.3, 1.0, 2.0, 2.5])
  all_t, mask = model._prepare_grid(times, None)
  rs = np.random.RandomState(1)
  z = rs.standard_normal((5000, all_t.shape[0] - 1, 1)).astype(npd)
  got = _np(model.sample_paths(times, normal_draws=z))[..., 0]
  p = ohw.Params(0.07, [0.5, 2.0], [0.01, 0.02, 0.015])
  want = ohw.sample_paths(p, all_t, R01, z[..., 0].astype(np.float64))
  np.testing.assert_allclose(got, want[:, mask], rtol=tol, atol=tol * 0.01)
  curve = np.array([0.0, 0.5, 1.0, 3.0])
  bonds, rates = model.sample_discount_curve_paths(times, curve,
                                                   normal_draws=z)
  assert tuple(bonds.shape) == (5000, 4, 4, 1)
  r = want[:, mask]
  want_b = ohw.bond_price(p, r[:, None, :], times[None, None, :],
                          times[None, None, :] + curve[None, :, None], R01,
                          R01)
  np.testing.assert_allclose(_np(bonds)[..., 0], want_b, rtol=tol)


def test_moments_and_bond_price_golden():
  a, s = 0.1, 0.01
  model = tqf.models.hull_white.HullWhiteModel1F(a, s, _flat,
                                                 dtype=torch.float64)
  r = _np(model.sample_paths([0.5, 1.0], num_samples=400000,
                             random_type=RandomType.STATELESS_ANTITHETIC,
                             seed=[1, 2]))[:, 1, 0]
  mean = 0.01 + s * s / 2 / a**2 * (1 - np.exp(-a))**2
  var = s * s / 2 / a * (1 - np.exp(-2 * a))
  assert abs(r.mean() - mean) < 4 * np.sqrt(var / r.size)
  np.testing.assert_allclose(r.var(), var, rtol=1e-2)
  bp = _np(model.discount_bond_price([[0.011], [0.01]], [1.0, 2.0],
                                     [2.0, 3.5]))
  assert bp.shape == (2, 1)
  np.testing.assert_allclose(bp[:, 0], [0.98906753, 0.98495442], rtol=1e-6)


def _swaptions():
  ex1 = [1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5]
  ex2 = [2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0, 5.0]
  # floating-leg start rows of bermudan_swaption_test.py:37-60: the row of
  # exercise date i starts at base[i] (5nc1) or base[i + 2] (5nc2),
  # capped at 5.0
  base = [1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0]
  s1 = [[base[min(i + j, 8)] for j in range(8)] for i in range(8)]
  s2 = [[base[min(i + j, 8)] for j in range(8)] for i in range(2, 10)]
  st = np.array([s1, s2])
  end = np.clip(st + 0.5, 0.0, 5.0)
  dcf = end - st
  return np.array([ex1, ex2]), st, end, dcf, 0.011 * np.ones_like(end)


def _price(ex, st, end, dcf, cpn, n, dtype=torch.float64):
  return _np(tqf.models.hull_white.bermudan_swaption_price(
      exercise_times=ex, floating_leg_start_times=st,
      floating_leg_end_times=end, fixed_leg_payment_times=end,
      floating_leg_daycount_fractions=dcf, fixed_leg_daycount_fractions=dcf,
      fixed_leg_coupon=cpn, reference_rate_fn=_flat, notional=100.0,
      mean_reversion=0.03, volatility=0.01, num_samples=n, time_step=0.1,
      random_type=RandomType.STATELESS_ANTITHETIC, seed=[0, 0],
      dtype=dtype))


# The reference's tolerances (bermudan_swaption_test.py) sit at or below
# the Monte Carlo error of its own 10^4-5*10^4 paths; these tests use 2^20
# paths, where the standard errors are about 0.0015 (5nc1 / 5nc2), 0.0025
# (one-date, vs Jamshidian) and 0.0005 / 0.001 (the QuantLib pair).
N_BIG = 2**20


def test_bermudan_reference_values():
  ex, st, end, dcf, cpn = _swaptions()
  one = _price(ex[0], st[0], end[0], dcf[0], cpn[0], N_BIG)
  assert one.shape == ()
  np.testing.assert_allclose(one, 1.8892, rtol=1e-2, atol=1e-2)
  both = _price(ex, st, end, dcf, cpn, N_BIG)
  assert both.shape == (2,)
  np.testing.assert_allclose(both, [1.8892, 1.6633], rtol=5e-3, atol=5e-3)


def test_bermudan_one_date_is_european():
  ex, st, end, dcf, cpn = _swaptions()
  berm = _price([ex[0][0]], st[0][0], end[0][0], dcf[0][0], cpn[0][0],
                N_BIG)
  euro = 100.0 * ohw.european_payer_swaption(0.03, 0.01, ex[0][0], end[0][0],
                                             dcf[0][0], cpn[0][0], R01)
  # 3 standard errors
  np.testing.assert_allclose(berm, euro, rtol=0, atol=0.0075)


def test_bermudan_quantlib_values():
  # bermudan_swaption_test.py:243-293: QuantLib finite-difference prices
  ex = [[0.498630137, 1.002739726], [1.002739726, 2.002739726]]
  st = [[[0.498630137, 1.002739726, 1.498630137, 2.002739726, 2.498630137],
         [1.002739726, 1.498630137, 2.002739726, 2.498630137, 3.002739726]],
        [[1.002739726, 1.498630137, 2.002739726, 2.498630137, 3.002739726],
         [2.002739726, 2.498630137, 3.002739726, 3.002739726, 3.002739726]]]
  end = [[[1.002739726, 1.498630137, 2.002739726, 2.498630137, 3.002739726],
          [1.498630137, 2.002739726, 2.498630137, 3.002739726, 3.002739726]],
         [[1.498630137, 2.002739726, 2.498630137, 3.002739726, 3.002739726],
          [2.498630137, 3.002739726, 3.002739726, 3.002739726, 3.002739726]]]
  dcf = np.array(end) - np.array(st)
  got = _np(tqf.models.hull_white.bermudan_swaption_price(
      exercise_times=ex, floating_leg_start_times=st,
      floating_leg_end_times=end, fixed_leg_payment_times=end,
      floating_leg_daycount_fractions=dcf, fixed_leg_daycount_fractions=dcf,
      fixed_leg_coupon=0.008 * np.ones_like(dcf), reference_rate_fn=_flat,
      notional=100.0, mean_reversion=0.03, volatility=0.01,
      num_samples=N_BIG, time_step=0.1,
      random_type=RandomType.STATELESS_ANTITHETIC, seed=[0, 0],
      dtype=torch.float64))
  # 3 standard errors at 2^20 paths
  np.testing.assert_allclose(got, [1.1016533104878343, 1.0586289990592714],
                             rtol=0, atol=0.003)


def test_bermudan_pipeline_matches_oracle_on_same_paths():
  ex, st, end, dcf, cpn = _swaptions()
  plan = sw._plan(ex, end, 0.1)
  model = tqf.models.hull_white.HullWhiteModel1F(0.03, 0.01, _flat,
                                                 dtype=torch.float64)
  all_t, mask = model._prepare_grid(plan['sim_times'], None)
  rs = np.random.RandomState(3)
  z = rs.standard_normal((20000, all_t.shape[0] - 1, 1))
  curve, rates = model.sample_discount_curve_paths(
      plan['sim_times'], plan['curve_times'], normal_draws=z)
  bonds = sw.bonds_from_curve_paths(plan, curve, end)
  # the pricer's own bond computation (exercise dates only) agrees with
  # the full discount-curve tensor
  np.testing.assert_allclose(_np(sw._exercise_bonds(model, plan, rates, end)),
                             _np(bonds), rtol=1e-14)
  got = _np(sw._price_on_paths(plan, bonds, rates, end, dcf, cpn,
                               tqf.models.longstaff_schwartz
                               .make_polynomial_basis(2), torch.float64))
  # the oracle: same draws, oracle paths, bonds, payoffs and LSM
  p = ohw.Params(0.03, [], [0.01])
  r = ohw.sample_paths(p, all_t, R01, z[..., 0])[:, mask]
  k, m = plan['sim_times'], plan['curve_times']
  pb = ohw.bond_price(p, r[:, None, :], k[None, None, :],
                      k[None, None, :] + m[None, :, None], R01, R01)
  dt = np.concatenate([[0.0], np.diff(k)])
  disc = np.cumprod(np.exp(-r * dt), axis=1)
  u = plan['unique_ex']
  ui = np.searchsorted(k, u)
  table = np.zeros((r.shape[0], 2, u.shape[0]))
  for b in range(2):
    for e in range(ex.shape[1]):
      ke = np.searchsorted(k, ex[b, e])
      taus = end[b, e] - ex[b, e]
      pj = pb[:, np.searchsorted(m, taus), ke]
      swap = (1 - pj[:, -1]) - (cpn[b, e] * dcf[b, e] * pj).sum(-1)
      table[:, b, np.searchsorted(u, ex[b, e])] = np.maximum(swap, 0)
  want = olsm.least_square_mc(r[:, ui, None], list(range(u.shape[0])),
                              lambda x, i: table[:, :, i],
                              olsm.make_polynomial_basis(2),
                              discount_factors=disc[:, None, ui])
  # the kernels solve the standardised normal equations, the oracle uses
  # pinv: the same exercise decisions up to rounding of the regression
  np.testing.assert_allclose(got, want, rtol=2e-6)


def test_time_dependent_mean_reversion_takes_the_euler_route():
  # hull_white_test.py:450-463: a mean reversion with jumps is sampled with
  # Euler (drift f'(0,t) + a f(0,t) + sigma^2 (1 - e^{-2at}) / (2a) - a r).
  # With equal values on both sides of the jump the law is the constant-a
  # model's: moments at t = 1 within Monte Carlo and Euler error.
  mr = PiecewiseConstantFunc([0.5], [0.1, 0.1])
  model = tqf.models.hull_white.HullWhiteModel1F(mr, 0.01, _flat,
                                                 dtype=torch.float64)
  assert model._sample_with_generic
  with pytest.raises(ValueError):
    model.sample_paths([1.0], num_samples=10)
  r = _np(model.sample_paths([1.0], num_samples=200000, time_step=0.01,
                             random_type=RandomType.STATELESS_ANTITHETIC,
                             seed=[4, 2]))[:, 0, 0]
  a, s = 0.1, 0.01
  mean = 0.01 + s * s / 2 / a**2 * (1 - np.exp(-a))**2
  var = s * s / 2 / a * (1 - np.exp(-2 * a))
  assert abs(r.mean() - mean) < 4 * np.sqrt(var / r.size)
  np.testing.assert_allclose(r.var(), var, rtol=2e-2)
  # the drift at a point against the formula
  x = torch.full((3, 1), 0.02, dtype=torch.float64)
  d = _np(model._drift_fn(torch.tensor(0.7, dtype=torch.float64), x))
  want = 0.0 + a * 0.01 + s * s / 2 / a * (1 - np.exp(-2 * a * 0.7)) - a * 0.02
  np.testing.assert_allclose(d, want, rtol=1e-12)
