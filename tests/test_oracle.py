"""Pins the CPU oracle against the reference's golden vectors and against
closed forms / independent constructions (no GPU needed)."""
import numpy as np
import pytest
from scipy import special, stats

from conftest import COEF_TIMING_GRID, load_golden, right_riemann_drift
from oracle import closed_form, euler, gbm, grid, heston, lsm, normal, rng, sobol


# ---------------------------------------------------------------- Sobol ----

def test_sobol_random_ops_goldens():
  g = load_golden('sobol_golden.json')
  e = g['random_ops_dim2_first5']
  for dt in (np.float16, np.float32, np.float64):
    got = sobol.sample(e['dim'], e['num_results'], dtype=dt)
    np.testing.assert_allclose(got, e['expected'], rtol=1e-6)
  e = g['random_ops_dim5_31_as_set']
  got = sobol.sample(5, 31)
  assert sorted(map(tuple, got.tolist())) == sorted(map(tuple, e['expected']))
  e = g['random_ops_large_skip']
  np.testing.assert_allclose(sobol.sample(1, 3, skip=e['skip']), e['expected'])


def test_sobol_skip_consistency():
  a = sobol.sample(10, 50 + 17)
  b = sobol.sample(10, 50, skip=17)
  np.testing.assert_allclose(a[17:], b)


def test_sobol_qmc_goldens():
  g = load_golden('sobol_golden.json')
  np.testing.assert_allclose(sobol.sobol_sample(5, 29),
                             g['qmc_sobol_29x5']['expected'], rtol=1e-6)
  e = g['qmc_sequence_indices']
  np.testing.assert_allclose(
      sobol.sobol_sample(5, 29, sequence_indices=e['indices']), e['expected'])
  np.testing.assert_allclose(
      sobol.sobol_sample(6, 8, apply_tent_transform=True),
      g['qmc_tent_6x8']['expected'])
  for dt in (np.float32, np.float64):
    np.testing.assert_allclose(sobol.sobol_sample(6, 8, dtype=dt),
                               g['qmc_dtype_6x8']['expected'])
  for key in ('qmc_generating_matrices', 'qmc_generating_matrices_doc'):
    e = g[key]
    np.testing.assert_array_equal(
        sobol.sobol_generating_matrices(e['dim'], e['num_results'],
                                        e['num_digits']), e['expected'])
  e = g['digital_net_dtype_5x6']
  gm = sobol.sobol_generating_matrices(5, 6, 3)
  np.testing.assert_allclose(sobol.digital_net_sample(gm, 6, 3), e['expected'])


def test_sobol_gray_code_cross_check():
  # An independent Gray-code (Antonov-Saleev) generator visits natural index
  # g(n) = n ^ (n >> 1) at step n.
  g = sobol.gray_code_ints(40, 1000, nd=10)
  nat, nd = sobol.sample_int(40, 1023)
  assert nd == 10
  n = np.arange(1, 1001)
  np.testing.assert_array_equal(g, nat[(n ^ (n >> 1)) - 1])


def test_sobol_net_property_and_index_agreement():
  # every aligned 2^m block of a dimension is a permutation of j / 2^m
  ints, nd = sobol.sample_int(30, 2**10 - 1)
  full = np.concatenate([np.zeros((1, 30), np.uint64), ints])  # index 0..1023
  for m in (4, 7):
    for b in range(0, 1024, 2**m):
      blk = full[b:b + 2**m] >> np.uint64(nd - m)
      for d in range(30):
        assert np.array_equal(np.sort(blk[:, d]), np.arange(2**m))
  # random_ops index k equals qmc index k for k >= 1 (same digits)
  q = sobol.digital_net_sample_int(sobol.sobol_generating_matrices(30, 1024, 10),
                                   1024, 10)
  np.testing.assert_array_equal(q[1:], ints.astype(np.int64))


# ---------------------------------------------------------------- Philox ---

def test_philox_known_answers():
  kat = load_golden('philox_kat.json')
  for v in kat['vectors']:
    c = np.array([int(x, 16) for x in v['counter']], np.uint64)
    k = np.array([int(x, 16) for x in v['key']], np.uint64)
    got = rng.philox4x32_10(c, k)
    assert [int(x) for x in got] == [int(x, 16) for x in v['expected']]


def test_box_muller_distribution():
  z = rng.normals(100000, 4, seed=42)
  assert abs(z.mean()) < 4 / np.sqrt(z.size)
  assert abs(z.var() - 1) < 4 * np.sqrt(2 / z.size)
  assert stats.kstest(z.ravel(), 'norm').pvalue > 1e-3
  # columns independent
  c = np.corrcoef(z.T)
  assert np.max(np.abs(c - np.eye(4))) < 0.02
  # closed form on one counter
  w = rng.raw_words(1, 2, seed=5)
  u = (float(w[0, 0]) + 1) * 2.0**-32
  t = float(w[0, 1]) * 2.0**-32
  want = np.sqrt(-2 * np.log(u)) * np.array([np.cos(2 * np.pi * t),
                                             np.sin(2 * np.pi * t)])
  np.testing.assert_allclose(rng.normals(1, 2, seed=5)[0], want, rtol=1e-14)
  anti = rng.normals(8, 3, seed=1, antithetic=True)
  np.testing.assert_array_equal(anti[:4], -anti[4:])


def test_normal_transforms():
  u = np.array([0.1, 0.25, 0.5, 0.75, 0.9])
  np.testing.assert_allclose(normal.quasi_normals(u), special.ndtri(u),
                             rtol=1e-13, atol=1e-15)
  ints = np.array([0, 2**31, 2**32 - 1, 12345678], np.uint64)
  z = normal.shifted_sobol_normals(ints)
  n = (ints >> np.uint64(7)).astype(np.float64)
  np.testing.assert_allclose(z, special.ndtri((n + 0.5) / 2**25), rtol=1e-12)


# ---------------------------------------------------------------- grids ----

def test_grid_from_num_times():
  t, mask, idx = grid.prepare_grid([1.0], np.float64, num_time_steps=100)
  assert t.shape == (101,) and t[0] == 0 and t[-1] == 1.0
  np.testing.assert_allclose(np.diff(t), 0.01, rtol=1e-12)
  assert mask.sum() == 1 and mask[-1]
  t, mask, idx = grid.prepare_grid([0.25, 0.5, 0.75, 1.0], np.float64,
                                   num_time_steps=365)
  assert t.shape == (366,)
  np.testing.assert_allclose(t[idx], [0.25, 0.5, 0.75, 1.0])
  assert np.all(np.diff(t) > 0)


def test_grid_from_time_step_dedup():
  t, mask, idx = grid.prepare_grid([0.1, 0.5, 1.0], np.float64, time_step=0.1)
  assert np.all(np.diff(t) > 1e-10)
  np.testing.assert_allclose(t[idx], [0.1, 0.5, 1.0], atol=1e-12)
  assert mask.sum() == 3


# ---------------------------------------------------------------- Euler ----

def test_euler_gbm_product_form():
  rs = np.random.RandomState(0)
  n, steps = 64, 50
  z = rs.standard_normal((n, steps, 1))
  drift, vol = euler.gbm_fns(0.05, 0.2)
  out = euler.sample(1, drift, vol, [1.0], num_samples=n, initial_state=[100.],
                     normal_draws=z, num_time_steps=steps)
  t, _, _ = grid.prepare_grid([1.0], np.float64, num_time_steps=steps)
  dt = np.diff(t)
  prod = 100 * np.prod(1 + 0.05 * dt + 0.2 * np.sqrt(dt) * z[:, :, 0], axis=1)
  np.testing.assert_allclose(out[:, 0, 0], prod, rtol=1e-12)


def test_euler_wiener_is_cumsum_and_records():
  rs = np.random.RandomState(1)
  times = [0.1, 0.35, 1.0]
  drift = lambda t, x: np.zeros_like(x)
  vol = lambda t, x: np.ones(x.shape + (1,))
  t, mask, idx = grid.prepare_grid(times, np.float64, num_time_steps=20)
  z = rs.standard_normal((5, t.shape[0] - 1, 1))
  out = euler.sample(1, drift, vol, times, num_samples=5, normal_draws=z,
                     num_time_steps=20)
  w = np.concatenate([np.zeros((5, 1)),
                      np.cumsum(z[:, :, 0] * np.sqrt(np.diff(t)), 1)], 1)
  np.testing.assert_allclose(out[:, :, 0], w[:, idx], rtol=1e-12, atol=1e-15)


def test_euler_coefficients_are_evaluated_at_the_next_grid_time():
  # zero volatility, drift mu sqrt(t): the Euler value is exactly the right
  # Riemann sum of the drift (euler_sampling.py:519 reads times[i + 1]);
  # the left sum differs by ~1e-2 here, far outside the tolerance
  mu, x0 = 0.2, 0.1
  drift = lambda t, x: mu * np.sqrt(t) * np.ones_like(x)
  vol = lambda t, x: np.zeros(x.shape + (1,))
  n_steps = COEF_TIMING_GRID.shape[0] - 1
  out = euler.sample(1, drift, vol, [0.3, 1.0], num_samples=3,
                     initial_state=[x0], normal_draws=np.zeros((3, n_steps, 1)),
                     times_grid=COEF_TIMING_GRID)
  right, left = right_riemann_drift(mu, x0, COEF_TIMING_GRID)
  np.testing.assert_allclose(out[:, 0, 0], right[3], rtol=1e-14)
  np.testing.assert_allclose(out[:, 1, 0], right[-1], rtol=1e-14)
  assert abs(right[-1] - left[-1]) > 1e-2


def test_euler_time_dependent_process_mean():
  # euler_sampling_test.py:206-294: dX = mu sqrt(t) dt + (a t + b) dW has
  # E[X_t] = x0 + 2/3 mu t^1.5; antithetic draws, the reference's grid
  mu, a, b, x0 = 0.2, 0.4, 0.33, 0.1
  times = np.array([0.0, 0.1, 0.21, 0.32, 0.43, 0.55])
  drift = lambda t, x: mu * np.sqrt(t) * np.ones_like(x)
  vol = lambda t, x: (a * t + b) * np.ones(x.shape + (1,))
  t, _, _ = grid.prepare_grid(times, np.float64, time_step=0.01)
  assert t.shape[0] - 1 == 55
  z = np.random.RandomState(42).standard_normal((5000, 55, 1))
  z = np.concatenate([z, -z], 0)
  out = euler.sample(1, drift, vol, times, num_samples=10000,
                     initial_state=[x0], normal_draws=z, time_step=0.01)
  assert out.shape == (10000, 6, 1)
  np.testing.assert_allclose(out.mean(0)[:, 0],
                             x0 + 2.0 / 3.0 * mu * times**1.5, rtol=1e-2,
                             atol=1e-2)
  # with antithetic draws the mean is the deterministic right Riemann sum
  right, _ = right_riemann_drift(mu, x0, t)
  _, _, idx = grid.prepare_grid(times, np.float64, time_step=0.01)
  np.testing.assert_allclose(out.mean(0)[1:, 0], right[idx[1:] - 1],
                             rtol=1e-12)


# ---------------------------------------------------------------- Heston ---

def test_heston_coefficient_goldens():
  g = load_golden('heston_coefficients.json')
  d, v = heston.drift_vol(np.array([g['state']]), g['kappa'], g['theta'],
                          g['volvol'], g['rho'])
  np.testing.assert_allclose(d[0], g['drift'], rtol=1e-6)
  np.testing.assert_allclose(v[0], g['vol'], rtol=1e-6, atol=1e-8)
  e = g['expected_total_variance']
  got = heston.expected_total_variance(e['T'], e['v0'], e['kappa'], e['theta'])
  want = ((e['v0'] - e['theta']) * (1 - np.exp(-e['kappa'] * e['T'])) /
          e['kappa'] + e['theta'] * e['T'])
  np.testing.assert_allclose(got, want)


def test_heston_semi_analytic_reduces_to_black_scholes():
  # the price moves linearly in xi (the rho*xi term) as xi -> 0
  bs = closed_form.bs_price(100, 90, 1.0, 0.03, 0.2)
  e3 = heston.european_call(100, 90, 1.0, 0.03, 0.04, 2.0, 0.04, 1e-3, -0.7)
  e4 = heston.european_call(100, 90, 1.0, 0.03, 0.04, 2.0, 0.04, 1e-4, -0.7)
  e5 = heston.european_call(100, 90, 1.0, 0.03, 0.04, 2.0, 0.04, 1e-5, -0.7)
  np.testing.assert_allclose(e5, bs, rtol=3e-6)
  np.testing.assert_allclose((e3 - bs) / (e4 - bs), 10.0, rtol=0.02)
  # put-call parity holds for the integral with stochastic vol
  c = heston.european_call(100, 110, 0.5, 0.03, 0.04, 2.0, 0.05, 0.5, -0.7)
  assert 0 < c < 100


def test_heston_qe_matches_conditional_mean():
  rs = np.random.RandomState(2)
  n = 200000
  z = rs.standard_normal((n, 1, 2))
  out = heston.sample_paths_qe([0.5], np.log(100.), 0.04, 1.5, 0.05, 0.6,
                               -0.5, z, num_time_steps=1)
  # E[v_T | v_0] = theta + (v0 - theta) e^{-kappa T}
  want = 0.05 + (0.04 - 0.05) * np.exp(-1.5 * 0.5)
  v = out[:, -1, 1]
  assert abs(v.mean() - want) < 4 * v.std() / np.sqrt(n)


def test_heston_qe_small_volvol_is_gbm():
  rs = np.random.RandomState(3)
  n = 100000
  z = rs.standard_normal((n, 10, 2))
  out = heston.sample_paths_qe([1.0], np.log(100.), 0.04, 1.0, 0.04, 1e-3,
                               0.0, z, num_time_steps=10)
  s = np.exp(out[:, -1, 0])
  assert abs(s.mean() - 100.) < 4 * s.std() / np.sqrt(n)
  v = out[:, -1, 1]
  np.testing.assert_allclose(v, 0.04, rtol=2e-2)
  assert abs(v.mean() - 0.04) < 4 * v.std() / np.sqrt(n)


# strongly correlated, high vol-of-vol Heston: every rho / xi term of the
# QE log-spot update (heston_model.py:554-572) moves these prices
QE_CASE = dict(kappa=1.5, theta=0.04, xi=0.6, rho=-0.7, v0=0.04, s0=100.0,
               t=1.0, strikes=(70.0, 85.0, 100.0, 115.0, 130.0))


def _qe_terminal_spots(n_half, steps, seed, rho_spot_sign=1.0):
  """QE terminal spots with antithetic draws; `rho_spot_sign = -1` plants a
  sign error in the log-spot update's correlation terms only."""
  c = QE_CASE
  dt = c['t'] / steps
  z = np.random.RandomState(seed).standard_normal((n_half, steps, 2))
  z = np.concatenate([z, -z], 0)
  x = np.full(2 * n_half, np.log(c['s0']))
  v = np.full(2 * n_half, c['v0'])
  for i in range(steps):
    vn = heston.update_variance(c['kappa'], c['theta'], c['xi'], v, dt,
                                z[:, i, 0])
    x = heston.update_log_spot(c['kappa'], c['theta'], c['xi'],
                               rho_spot_sign * c['rho'], v, vn, x, dt,
                               z[:, i, 1])
    v = vn
  return np.exp(x)


def _qe_call_errors_in_se(spots):
  c = QE_CASE
  errs = []
  for k in c['strikes']:
    pay = np.maximum(spots - k, 0)
    want = heston.european_call(c['s0'], k, c['t'], 0.0, c['v0'], c['kappa'],
                                c['theta'], c['xi'], c['rho'])
    errs.append((pay.mean() - want) / (pay.std(ddof=1) / np.sqrt(spots.size)))
  return np.array(errs)


def test_heston_qe_log_spot_prices_match_semi_analytic_with_correlation():
  # the reference's own cross-check (heston_model_test.py:290-380), at
  # rho = -0.7 and xi = 0.6 where the QE spot update's correlation terms
  # carry the skew: the QE oracle's call prices over the smile and its
  # martingale mean agree with the independent Lewis integral
  spots = _qe_terminal_spots(2**16, 50, 11)
  errs = _qe_call_errors_in_se(spots)
  assert np.all(np.abs(errs) < 3), errs
  assert abs(spots.mean() - 100.0) < 3 * spots.std() / np.sqrt(spots.size)
  # a sign error in those terms (planted here, in the oracle's inputs) moves
  # the smile by tens of standard errors: the check discriminates
  bad = _qe_call_errors_in_se(_qe_terminal_spots(2**16, 50, 11, -1.0))
  assert np.max(np.abs(bad)) > 10, bad


# ---------------------------------------------------------------- GBM ------

def test_gbm_exact_single_step_and_moments():
  rs = np.random.RandomState(4)
  z = rs.standard_normal((100000, 2, 1))
  out = gbm.sample_paths([0.5, 1.0], 0.05, 0.2, 100.0, z)
  np.testing.assert_allclose(
      out[:, 0, 0], 100 * np.exp((0.05 - 0.02) * 0.5 + 0.2 * np.sqrt(.5) *
                                 z[:, 0, 0]), rtol=1e-13)
  lg = np.log(out[:, 1, 0] / 100)
  assert abs(lg.mean() - 0.03) < 3 * lg.std() / np.sqrt(lg.size)
  assert abs(lg.var() - 0.04) < 3 * 0.04 * np.sqrt(2 / (lg.size - 1))


def test_mv_gbm_correlation():
  rs = np.random.RandomState(5)
  corr = np.array([[1, .5], [.5, 1]])
  z = rs.standard_normal((100000, 1, 2))
  out = gbm.mv_sample_paths([1.0], [0.0, 0.0], [0.2, 0.3], corr, [1., 1.], z)
  lg = np.log(out[:, 0])
  assert abs(np.corrcoef(lg.T)[0, 1] - 0.5) < 0.01


# ---------------------------------------------------------------- LSM ------

def test_lsm_longstaff_schwartz_table():
  g = load_golden('lsm_ls2001.json')
  s = np.array(g['samples'])[..., None]
  df = np.exp(-np.cumsum(g['interest_rates']))
  tol = g['tolerance']
  basis = lsm.make_polynomial_basis(2)
  put = lsm.make_basket_put_payoff([1.1])
  for dt in (np.float32, np.float64):
    np.testing.assert_allclose(
        lsm.least_square_mc(s, [3], put, basis, discount_factors=[df[-1]],
                            dtype=dt), [g['european_put']], atol=tol, rtol=tol)
    np.testing.assert_allclose(
        lsm.least_square_mc(s, [1, 2, 3], put, basis, discount_factors=df,
                            dtype=dt), [g['american_put']], atol=tol, rtol=tol)
    np.testing.assert_allclose(
        lsm.least_square_mc(s, [1, 2, 3], put, basis, discount_factors=df,
                            num_calibration_samples=4, dtype=dt),
        [g['american_put_calib4']], atol=tol, rtol=tol)
  put2 = lsm.make_basket_put_payoff([1.1, 1.2])
  df2 = np.exp(-np.cumsum(g['batch_rates'], -1))[None]
  np.testing.assert_allclose(
      lsm.least_square_mc(s, [1, 2, 3], put2, basis, discount_factors=df2),
      g['batch_payoff'], atol=tol, rtol=tol)
  np.testing.assert_allclose(
      lsm.least_square_mc(np.stack([s, s + 0.1]), [1, 2, 3], put2, basis,
                          discount_factors=df2),
      g['batch_samples'], atol=tol, rtol=tol)


def test_lsm_duplicate_asset_invariant_and_single_date():
  g = load_golden('lsm_ls2001.json')
  s = np.array(g['samples'])[..., None]
  df = np.exp(-np.cumsum(g['interest_rates']))
  basis = lsm.make_polynomial_basis(10)
  put = lsm.make_basket_put_payoff([1.1, 1.2, 1.3])
  a = lsm.least_square_mc(np.concatenate([s, s], -1), [1, 2, 3], put, basis,
                          discount_factors=df)
  b = lsm.least_square_mc(s, [1, 2, 3], put, basis, discount_factors=df)
  np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-4)
  # one exercise date = discounted European mean
  e = lsm.least_square_mc(s, [3], lsm.make_basket_put_payoff([1.1]),
                          lsm.make_polynomial_basis(2),
                          discount_factors=[df[-1]])
  np.testing.assert_allclose(
      e, df[-1] * np.maximum(1.1 - s[:, 3, 0], 0).mean(), rtol=1e-12)


def test_polynomial_basis_order():
  x = np.array([[1.0], [2.0], [3.0], [4.0]])[..., None]  # [4, 1, 1]
  b = lsm.make_polynomial_basis(2)(x, 0)
  np.testing.assert_allclose(b[0], [[1, 1, 1, 1], [-1.5, -.5, .5, 1.5],
                                    [2.25, .25, .25, 2.25]])


# ---------------------------------------------------------------- closed ---

def test_black_scholes_parity_and_greeks():
  c = closed_form.bs_price(100, 95, 0.7, 0.03, 0.25)
  p = closed_form.bs_price(100, 95, 0.7, 0.03, 0.25, is_call=False)
  np.testing.assert_allclose(c - p, 100 - 95 * np.exp(-0.03 * 0.7))
  h = 1e-5
  fd = (closed_form.bs_price(100 + h, 95, .7, .03, .25) -
        closed_form.bs_price(100 - h, 95, .7, .03, .25)) / (2 * h)
  np.testing.assert_allclose(closed_form.bs_delta(100, 95, .7, .03, .25), fd,
                             rtol=1e-7)
  fd = (closed_form.bs_price(100, 95, .7, .03, .25 + h) -
        closed_form.bs_price(100, 95, .7, .03, .25 - h)) / (2 * h)
  np.testing.assert_allclose(closed_form.bs_vega(100, 95, .7, .03, .25), fd,
                             rtol=1e-7)
  fd = (closed_form.bs_price(100, 95, .7, .03 + h, .25) -
        closed_form.bs_price(100, 95, .7, .03 - h, .25)) / (2 * h)
  np.testing.assert_allclose(closed_form.bs_rho(100, 95, .7, .03, .25), fd,
                             rtol=1e-6)


def test_basket_pathwise_greeks_match_finite_differences():
  rs = np.random.RandomState(6)
  d = 4
  a = rs.standard_normal((d, d))
  c = a @ a.T + np.eye(d)
  corr = c / np.sqrt(np.outer(np.diag(c), np.diag(c)))
  s0 = np.full(d, 100.0)
  sig = rs.uniform(0.15, 0.45, d)
  z = rs.standard_normal((20000, 8, d))
  p = gbm.basket_call_paths(s0, sig, corr, 0.02, 1.0, 100.0, z)
  h = 1e-6
  for i in range(d):
    e = np.zeros(d)
    e[i] = h * 100
    fd = (gbm.basket_call_price(s0 + e, sig, corr, .02, 1., 100., z) -
          gbm.basket_call_price(s0 - e, sig, corr, .02, 1., 100., z)) / (2 * e[i])
    np.testing.assert_allclose(p['delta'][:, i].mean(), fd, rtol=1e-5)
    e = np.zeros(d)
    e[i] = h
    fd = (gbm.basket_call_price(s0, sig + e, corr, .02, 1., 100., z) -
          gbm.basket_call_price(s0, sig - e, corr, .02, 1., 100., z)) / (2 * h)
    np.testing.assert_allclose(p['vega'][:, i].mean(), fd, rtol=1e-5)
  fd = (gbm.basket_call_price(s0, sig, corr, .02 + h, 1., 100., z) -
        gbm.basket_call_price(s0, sig, corr, .02 - h, 1., 100., z)) / (2 * h)
  np.testing.assert_allclose(p['rho'].mean(), fd, rtol=1e-5)


def test_mlp_oracle_gradient_matches_finite_differences():
  from oracle import mlp
  rs = np.random.RandomState(0)
  p = mlp.init_params(1) + 0.01 * rs.standard_normal(mlp.N_PARAMS)
  x = rs.standard_normal((40, 8))
  t = rs.standard_normal(40)
  w = (rs.uniform(size=40) > 0.3).astype(np.float64)
  loss, g = mlp.loss_and_grad(p, x, t, w, 1 / 40)
  idx = rs.choice(mlp.N_PARAMS, 60, replace=False)
  idx = np.concatenate([idx, [0, 383, 384, 447, 448, 4543, 4544, 4608, 4672]])
  h = 1e-6
  for i in idx:
    e = np.zeros_like(p)
    e[i] = h
    fd = (mlp.loss_and_grad(p + e, x, t, w, 1 / 40)[0] -
          mlp.loss_and_grad(p - e, x, t, w, 1 / 40)[0]) / (2 * h)
    np.testing.assert_allclose(g[i], fd, rtol=1e-5, atol=1e-9)


def test_local_vol_table_is_exact_on_bilinear_surfaces():
  from oracle import local_vol
  times = np.array([0.0, 0.5, 1.5])
  x = np.linspace(np.log(50.), np.log(200.), 7)
  a, b, c = 0.2, -0.03, 0.05
  vols = a + b * (x[None, :] - x[0]) + c * times[:, None]
  for t in (0.0, 0.2, 0.5, 1.0, 1.5):
    xs = np.linspace(x[0], x[-1], 11)
    got = local_vol.table_sigma(times, x[0], x[1] - x[0], vols, t, xs)
    np.testing.assert_allclose(got, a + b * (xs - x[0]) + c * t, rtol=1e-13)
  # flat outside the grid
  got = local_vol.table_sigma(times, x[0], x[1] - x[0], vols, 3.0,
                              np.array([x[0] - 1, x[-1] + 1]))
  np.testing.assert_allclose(got, [vols[-1, 0], vols[-1, -1]], rtol=1e-13)



def test_halton_oracle_known_values_and_radical_inverse():
  from fractions import Fraction
  from oracle import halton
  # halton_test.py:30-37
  want = np.array([[1. / 2, 1. / 3], [1. / 4, 2. / 3], [3. / 4, 1. / 9],
                   [1. / 8, 4. / 9], [5. / 8, 7. / 9]])
  np.testing.assert_allclose(halton.sample(2, num_results=5), want, rtol=1e-6)
  p = halton.primes(1000)
  assert list(p[:10]) == [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]
  assert p[-1] == 7919  # the 1000th prime
  # radical inverse with exact fractions, float64 truncation irrelevant here
  idx = np.array([0, 1, 7, 100, 12345, 999999])
  got = halton.sample(6, sequence_indices=idx, dtype=np.float64)
  for i, k in enumerate(idx):
    for d, b in enumerate(p[:6]):
      n, f, w = int(k) + 1, Fraction(0), Fraction(1)
      while n:
        w /= int(b)
        f += (n % int(b)) * w
        n //= int(b)
      assert abs(got[i, d] - float(f)) <= 1e-15
  np.testing.assert_array_equal(
      halton.sample(3, num_results=50),
      halton.sample(3, sequence_indices=np.arange(50)))
  # identity permutations and no correction leave the sequence unchanged
  bases = p[:3]
  ident = np.concatenate([np.arange(b) for b in bases])
  perms = np.tile(ident, (24, 1))
  np.testing.assert_array_equal(
      halton.sample(3, num_results=40, perms=perms),
      halton.sample(3, num_results=40))


def test_heston_ft_tangent_greeks_match_central_differences():
  # forward tangents through the full-truncation step loop against central
  # differences of the price on the same draws (Feller condition kept so
  # the truncation kinks do not dominate)
  rs = np.random.RandomState(8)
  n, steps = 20000, 40
  z = rs.standard_normal((n, steps, 2))
  t = np.linspace(0.0, 1.0, steps + 1)
  keep = np.zeros(steps + 1, bool)
  keep[-1] = True
  p = dict(x0=np.log(100.0), v0=0.04, kappa=2.0, theta=0.05, xi=0.3,
           rho=-0.6, r=0.03)
  mean, se = heston.ft_euler_call_greeks(
      p['x0'], p['v0'], p['kappa'], p['theta'], p['xi'], p['rho'], p['r'],
      100.0, t, z)

  def price(**kw):
    q = dict(p, **kw)
    pr, _ = heston.ft_euler_calls(q['x0'], q['v0'], q['kappa'], q['theta'],
                                  q['xi'], q['rho'], q['r'], [100.0], t, keep,
                                  z)
    return pr[-1, 0]
  np.testing.assert_allclose(mean[0], price(), rtol=1e-12)
  for k, (name, h) in enumerate([('s0', 1e-4), ('v0', 1e-7), ('r', 1e-7),
                                 ('kappa', 1e-6), ('theta', 1e-7),
                                 ('xi', 1e-6), ('rho', 1e-6)]):
    if name == 's0':
      fd = (price(x0=np.log(100.0 + h)) - price(x0=np.log(100.0 - h))) / (2 * h)
    else:
      fd = (price(**{name: p[name] + h}) - price(**{name: p[name] - h})) / (
          2 * h)
    assert abs(mean[k + 1] - fd) <= 1e-4 * abs(fd) + 0.01 * se[k + 1], (
        name, mean[k + 1], fd)
