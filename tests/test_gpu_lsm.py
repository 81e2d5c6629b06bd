"""GPU Longstaff-Schwartz against the reference's table values and the
CPU oracle (numpy pinv) on simulated paths."""
import numpy as np
import pytest
import torch

from conftest import load_golden
from oracle import closed_form, gbm as ogbm
from oracle import lsm as olsm

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
  import tqf
  lsm = tqf.models.longstaff_schwartz


def _np(t):
  return t.detach().cpu().numpy()


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_lsm_reference_table(dtype):
  g = load_golden('lsm_ls2001.json')
  s = np.array(g['samples'])[..., None]
  df = np.exp(-np.cumsum(g['interest_rates']))
  tol = g['tolerance']
  basis = lsm.make_polynomial_basis(2)
  put = lsm.make_basket_put_payoff([1.1], dtype=dtype)
  np.testing.assert_allclose(
      _np(lsm.least_square_mc(s, [3], put, basis, discount_factors=[df[-1]],
                              dtype=dtype)), [g['european_put']], atol=tol,
      rtol=tol)
  np.testing.assert_allclose(
      _np(lsm.least_square_mc(s, [1, 2, 3], put, basis, discount_factors=df,
                              dtype=dtype)), [g['american_put']], atol=tol,
      rtol=tol)
  np.testing.assert_allclose(
      _np(lsm.least_square_mc(s, [1, 2, 3], put, basis, discount_factors=df,
                              num_calibration_samples=4, dtype=dtype)),
      [g['american_put_calib4']], atol=tol, rtol=tol)
  put2 = lsm.make_basket_put_payoff([1.1, 1.2], dtype=dtype)
  df2 = np.exp(-np.cumsum(g['batch_rates'], -1))[None]
  np.testing.assert_allclose(
      _np(lsm.least_square_mc(s, [1, 2, 3], put2, basis, discount_factors=df2,
                              dtype=dtype)), g['batch_payoff'], atol=tol,
      rtol=tol)
  np.testing.assert_allclose(
      _np(lsm.least_square_mc(np.stack([s, s + 0.1]), [1, 2, 3], put2, basis,
                              discount_factors=df2, dtype=dtype)),
      g['batch_samples'], atol=tol, rtol=tol)


def test_lsm_duplicate_assets_equal_one_asset():
  g = load_golden('lsm_ls2001.json')
  s = np.array(g['samples'])[..., None]
  df = np.exp(-np.cumsum(g['interest_rates']))
  basis = lsm.make_polynomial_basis(10)
  put = lsm.make_basket_put_payoff([1.1, 1.2, 1.3], dtype=np.float64)
  a = lsm.least_square_mc(np.concatenate([s, s], -1), [1, 2, 3], put, basis,
                          discount_factors=df, dtype=np.float64)
  b = lsm.least_square_mc(s, [1, 2, 3], put, basis, discount_factors=df,
                          dtype=np.float64)
  np.testing.assert_allclose(_np(a), _np(b), rtol=1e-4, atol=1e-4)
  assert a.shape == (3,)


# (assets, degree) pairs cover the one-asset register moments and the MFMA
# moment product with row groups of 0, 1 and 2 dimensions and padded tiles:
# (1, 3) registers; (2, 2) 5 x 8 of 16 x 16; (3, 2) 25 x 8 of 32 x 16;
# (2, 5) 11 x 17 of 16 x 32; (3, 1) 3 x 13 of 16 x 16; (4, 2) 25 x 34 of
# 32 x 48
@pytest.mark.parametrize('dim,degree,kind', [(1, 3, 'put'), (2, 2, 'put'),
                                             (3, 2, 'max_call'),
                                             (2, 5, 'put'), (3, 1, 'max_call'),
                                             (4, 2, 'max_call')])
def test_lsm_matches_oracle_on_simulated_paths(dim, degree, kind):
  rs = np.random.RandomState(dim)
  n, times = 20000, np.linspace(0.25, 1.0, 4)
  corr = np.full((dim, dim), 0.3) + 0.7 * np.eye(dim)
  z = rs.standard_normal((n, times.shape[0], dim))
  paths = ogbm.mv_sample_paths(times, [0.02] * dim, [0.25] * dim, corr,
                               [100.0] * dim, z)
  paths = np.concatenate([np.full((n, 1, dim), 100.0), paths], 1)
  df = np.exp(-0.05 * times)
  if kind == 'put':
    pay_o = olsm.make_basket_put_payoff([100.0])
    pay_t = lsm.make_basket_put_payoff([100.0])
  else:
    pay_o = olsm.make_basket_call_max_payoff([100.0])
    pay_t = lsm.make_max_call_payoff([100.0])
  for calib in (None, n // 2):
    # the kernels solve the standardised normal equations exactly; the
    # oracle does the same with numpy pinv on the standardised basis
    want = olsm.least_square_mc(paths, [1, 2, 3, 4], pay_o,
                                olsm.make_polynomial_basis(degree, True),
                                discount_factors=df,
                                num_calibration_samples=calib)
    got = lsm.least_square_mc(paths, [1, 2, 3, 4], pay_t,
                              lsm.make_polynomial_basis(degree),
                              discount_factors=df,
                              num_calibration_samples=calib,
                              dtype=np.float64)
    # same exercise decisions up to rounding of the regression: the price
    # agrees to far below its Monte Carlo error
    np.testing.assert_allclose(_np(got), want, rtol=2e-6)


def test_bermudan_max_call_five_assets_matches_oracle():
  # the learned-LSM workload's shape: 5 correlated GBM assets (rho 0.3,
  # r 0.05, q 0.1), 9 exercise dates of 20 exact steps, max-call K = 100,
  # make_polynomial_basis(2): K = 243 columns (the 32 x 160 moment split,
  # the 243-column Cholesky and the 5-asset dense update), 2^15
  # calibration + 2^15 pricing paths
  print('lsm 5-asset max-call vs oracle', flush=True)
  from tqf.engine import mlp_lsm
  paths, df = mlp_lsm.bermudan_max_call_paths(2**16)
  assert paths.shape == (2**16, 10, 5)
  ex = list(range(1, 10))
  values = torch.empty(2**16, dtype=torch.float64, device=paths.device)
  got, se, ranks = lsm._native_lsm(
      paths, ex, (lsm.PAYOFF_MAX_CALL, np.array([100.0])), 2,
      lsm._padded_discounts(df, 9, torch.float64), 2**15, torch.float64,
      paths.device, return_details=True, path_values=values)
  assert int(ranks.min()) == 243
  p = _np(paths)
  want, want_pv = olsm.least_square_mc(
      p, ex, olsm.make_basket_call_max_payoff([100.0]),
      olsm.make_polynomial_basis(2, True), discount_factors=df,
      num_calibration_samples=2**15, return_path_values=True)
  np.testing.assert_allclose(_np(got), want, rtol=2e-6)
  # path by path: the same exercise decisions up to rounding of the fit
  pv = df[0] * _np(values)[2**15:]
  agree = np.isclose(pv, want_pv[:, 0], rtol=1e-9, atol=1e-12)
  assert agree.mean() > 0.999, agree.mean()
  # the public entry point takes the same kernels
  pub = lsm.least_square_mc(paths, ex, lsm.make_max_call_payoff([100.0]),
                            lsm.make_polynomial_basis(2),
                            discount_factors=df,
                            num_calibration_samples=2**15)
  np.testing.assert_allclose(_np(pub), _np(got), rtol=1e-12)
  # the reference's own semantics (pinv of the raw 243-column Gram with
  # TF's cutoff) discard most directions at this conditioning and exercise
  # far worse: a much lower lower-bound price on the same paths
  raw = olsm.least_square_mc(p, ex, olsm.make_basket_call_max_payoff([100.0]),
                             olsm.make_polynomial_basis(2),
                             discount_factors=df,
                             num_calibration_samples=2**15)
  print('5-asset max-call: standardised %.4f, raw-basis pinv %.4f' %
        (want[0], raw[0]), flush=True)
  assert raw[0] < want[0]


def test_generic_torch_lsm_matches_oracle_and_native():
  rs = np.random.RandomState(9)
  n, times = 5000, np.linspace(0.5, 2.0, 4)
  z = rs.standard_normal((n, 4, 1))
  paths = ogbm.sample_paths(times, 0.03, 0.3, 100.0, z)
  paths = np.concatenate([np.full((n, 1, 1), 100.0), paths], 1)
  df = np.exp(-0.03 * times)
  put = lsm.make_basket_put_payoff([100.0])
  basis = lsm.make_polynomial_basis(2)
  plain_basis = lambda p, t: basis(p, t)
  plain_put = lambda p, t: put(p, t)
  generic = lsm.least_square_mc(paths, [1, 2, 3, 4], plain_put, plain_basis,
                                discount_factors=df, dtype=np.float64)
  want = olsm.least_square_mc(paths, [1, 2, 3, 4],
                              olsm.make_basket_put_payoff([100.0]),
                              olsm.make_polynomial_basis(2),
                              discount_factors=df)
  np.testing.assert_allclose(_np(generic), want, rtol=1e-9)
  native = lsm.least_square_mc(paths, [1, 2, 3, 4], put, basis,
                               discount_factors=df, dtype=np.float64)
  np.testing.assert_allclose(_np(generic), _np(native), rtol=1e-9)


def test_reference_pinv_cutoff_loses_the_fit_on_ill_conditioned_bases():
  # Documented deviation: on a 3-asset degree-2 max-call the reference's
  # pinv of the raw Gram truncates directions and prices lower; the kernel
  # solves the least-squares problem (standardised basis) exactly.
  dim, n = 3, 20000
  rs = np.random.RandomState(dim)
  times = np.linspace(0.25, 1.0, 4)
  corr = np.full((dim, dim), 0.3) + 0.7 * np.eye(dim)
  z = rs.standard_normal((n, 4, dim))
  paths = ogbm.mv_sample_paths(times, [0.02] * dim, [0.25] * dim, corr,
                               [100.0] * dim, z)
  paths = np.concatenate([np.full((n, 1, dim), 100.0), paths], 1)
  df = np.exp(-0.05 * times)
  pay_o = olsm.make_basket_call_max_payoff([100.0])
  raw = olsm.least_square_mc(paths, [1, 2, 3, 4], pay_o,
                             olsm.make_polynomial_basis(2),
                             discount_factors=df)
  exact = olsm.least_square_mc(paths, [1, 2, 3, 4], pay_o,
                               olsm.make_polynomial_basis(2, True),
                               discount_factors=df)
  got = lsm.least_square_mc(paths, [1, 2, 3, 4],
                            lsm.make_max_call_payoff([100.0]),
                            lsm.make_polynomial_basis(2),
                            discount_factors=df, dtype=np.float64)
  np.testing.assert_allclose(_np(got), exact, rtol=2e-6)
  assert raw[0] < exact[0]


def _crr_american(s0, k, t, r, sigma, call, steps=4000):
  """Cox-Ross-Rubinstein binomial tree with early exercise (independent
  check of the American value)."""
  dt = t / steps
  u = np.exp(sigma * np.sqrt(dt))
  d = 1 / u
  p = (np.exp(r * dt) - d) / (u - d)
  disc = np.exp(-r * dt)
  j = np.arange(steps + 1)
  s = s0 * u**(steps - j) * d**j
  v = np.maximum(s - k, 0) if call else np.maximum(k - s, 0)
  for n in range(steps - 1, -1, -1):
    s = s[:-1] / u
    v = disc * (p * v[:-1] + (1 - p) * v[1:])
    v = np.maximum(v, np.maximum(s - k, 0) if call else np.maximum(k - s, 0))
  return v[0]


def test_bs_lsm_price_american_options():
  from tqf.experimental import american_option
  spots = np.array([100.0, 90.0, 100.0])
  strikes = np.array([110.0, 100.0, 95.0])
  expiries = np.array([1.0, 0.5, 1.0])
  vols = np.array([0.25, 0.3, 0.2])
  r = 0.05
  df = np.exp(-r * expiries)
  call = np.array([False, False, True])
  got = _np(american_option.bs_lsm_price(
      spots, expiries, strikes, vols, df, num_samples=100000,
      num_exercise_times=50, is_call_option=call, seed=(1, 2)))
  for i in range(3):
    want = _crr_american(spots[i], strikes[i], expiries[i], r, vols[i],
                         call[i])
    # LSM on 50 dates with 1e5 paths: low-biased by its sub-optimal rule
    # and by the discrete exercise grid, plus Monte Carlo noise
    assert abs(got[i] - want) <= 0.015 * want + 0.02, (i, got[i], want)
  # no early exercise for a call without dividends: Black-Scholes
  bs = closed_form.bs_price(spots[2], strikes[2], 1.0, r, 0.2)
  assert abs(got[2] - bs) <= 0.015 * bs


def test_bs_lsm_price_batch_on_streams_equals_one_option_at_a_time():
  # the options of a batch run their regressions on side streams (four,
  # so five options reuse one): each price equals that option priced on
  # its own, exactly, and a second batch call repeats the first
  from tqf.experimental import american_option
  spots = [90.0, 100.0, 110.0, 100.0, 95.0]
  expiries = [1.0, 1.0, 0.5, 2.0, 1.5]
  vols = [0.2, 0.25, 0.3, 0.2, 0.35]
  calls = [False, False, True, False, True]
  dfs = [float(np.exp(-0.05 * e)) for e in expiries]
  kw = dict(num_samples=50000, num_exercise_times=30, seed=(5, 6))
  batch = _np(american_option.bs_lsm_price(
      spots, expiries, [100.0] * 5, vols, dfs, is_call_option=calls, **kw))
  again = _np(american_option.bs_lsm_price(
      spots, expiries, [100.0] * 5, vols, dfs, is_call_option=calls, **kw))
  np.testing.assert_array_equal(batch, again)
  for i in range(5):
    one = _np(american_option.bs_lsm_price(
        [spots[i]], [expiries[i]], [100.0], [vols[i]], [dfs[i]],
        is_call_option=[calls[i]], **kw))[0]
    assert batch[i] == one, (i, batch[i], one)


def test_bs_lsm_price_matches_oracle_on_the_same_paths():
  # the GPU LSM (basis on ln S, payoff on S) against the oracle's
  # restatement of least_square_mc on the engine's own paths
  import tqf
  from tqf.experimental import american_option
  from tqf.math.random.multivariate_normal import RandomType
  n, m, s0, k, t, sig, r = 20000, 20, 100.0, 105.0, 0.75, 0.3, 0.04
  got = float(_np(american_option.bs_lsm_price(
      [s0], [t], [k], [sig], [np.exp(-r * t)], num_samples=n,
      num_exercise_times=m, is_call_option=False, seed=(3, 4)))[0])
  times = np.linspace(0.0, 1.0, m)
  unit = _np(tqf.models.GeometricBrownianMotion(
      0.0, 1.0, dtype=torch.float64).sample_paths(
          times, initial_state=1.0, num_samples=n, seed=(3, 4),
          random_type=RandomType.STATELESS_ANTITHETIC))[..., 0]
  w = np.log(unit) + 0.5 * times
  log_s = sig * np.sqrt(t) * w + (r - 0.5 * sig**2) * times * t
  kk = k / s0

  def payoff(x, ti):
    return np.maximum(kk - np.exp(x[:, ti, 0]), 0)[:, None]
  want = olsm.least_square_mc(log_s[..., None], list(range(m)), payoff,
                              olsm.make_polynomial_basis(2, True),
                              discount_factors=np.exp(-r * times * t))[0]
  # same exercise decisions up to rounding of the regression
  np.testing.assert_allclose(got, want * s0, rtol=2e-6)


def test_table_payoff_and_path_discounts_native_matches_torch_and_oracle():
  # exercise values given as a table and discount factors that differ by
  # path (the reference's rank-3 discount_factors, lsm.py:239-256): the
  # native kernels against the torch restatement and the oracle
  rs = np.random.RandomState(12)
  n, t = 30000, 6
  z = rs.standard_normal((n, t, 1))
  times = np.linspace(0.2, 1.2, t)
  paths = ogbm.sample_paths(times, 0.02, 0.3, 1.0, z)  # [N, T, 1]
  rate = 0.03 + 0.01 * rs.standard_normal((n, t))
  dfs = np.exp(-np.cumsum(rate * 0.2, axis=1))  # [N, T]
  strikes = np.array([1.0, 1.1])
  table = np.maximum(strikes[None, :, None] - paths[:, None, :, 0], 0)
  table_t = torch.as_tensor(table, device='cuda')
  basis = lsm.make_polynomial_basis(2)
  native = lsm.least_square_mc(paths, list(range(t)),
                               lsm.make_table_payoff(table_t), basis,
                               discount_factors=dfs[:, None, :],
                               dtype=np.float64)
  generic = lsm.least_square_mc(paths, list(range(t)),
                                lambda x, i: table_t[:, :, i],
                                lambda x, i: basis(x, i),
                                discount_factors=dfs[:, None, :],
                                dtype=np.float64)
  want = olsm.least_square_mc(paths, list(range(t)),
                              lambda x, i: table[:, :, i],
                              olsm.make_polynomial_basis(2, True),
                              discount_factors=dfs[:, None, :])
  np.testing.assert_allclose(_np(native), want, rtol=2e-6)
  np.testing.assert_allclose(_np(generic), want, rtol=2e-6)
  assert native.shape == (2,)


@pytest.mark.gpu
def test_total_degree_table_takes_the_dense_gram_and_basis_order_update():
  """A total-degree exponent table (5 assets, degree 2: 21 monomials) is
  cheaper as dense Gram tiles than as moments of the 5^5 box, and cheaper as
  a basis-order sum than as the 3^5 dense coefficient tensor, so it runs the
  other Gram and update kernels than the tensor grids of the other tests:
  its price matches the oracle with the same monomials."""
  import itertools
  rs = np.random.RandomState(21)
  n, d = 20000, 5
  times = np.linspace(0.25, 1.0, 4)
  z = rs.standard_normal((n, 4, d))
  paths = np.concatenate(
      [ogbm.sample_paths(times, 0.02, 0.25 + 0.02 * j, 100.0, z[..., j:j + 1])
       for j in range(d)], -1)
  paths = np.concatenate([np.full((n, 1, d), 100.0), paths], 1)
  df = np.exp(-0.02 * times)
  table = np.array([e for e in itertools.product(range(3), repeat=d)
                    if sum(e) <= 2], np.uint8)
  assert table.shape == (21, d)

  def basis(sample_paths, time_index):
    sl = np.asarray(sample_paths)[None][:, :, time_index, :]
    c = sl - sl.mean(axis=1, keepdims=True)
    out = np.prod(c[:, :, None, :]**table[None, None].astype(float), axis=-1)
    return np.transpose(out, (0, 2, 1))

  want = olsm.least_square_mc(paths, [1, 2, 3, 4],
                              olsm.make_basket_put_payoff([100.0]), basis,
                              discount_factors=df)
  from tqf import _native
  dev = _native.device()
  x = torch.as_tensor(paths, dtype=torch.float64, device=dev)
  pspec = lsm.make_basket_put_payoff([100.0])._tqf_payoff
  got = lsm._native_lsm(x, [1, 2, 3, 4], pspec, 2,
                        lsm._padded_discounts(df, 4, torch.float64), None,
                        torch.float64, dev, exponents=table)
  np.testing.assert_allclose(_np(got), want, rtol=2e-6)


@pytest.mark.parametrize('dim', [1, 2, 5])
def test_log_price_paths_equal_the_log_basis_on_prices(dim):
  # basis_log = 2 (paths hold ln S: basis on them, payoff on exp, the
  # one-asset in-the-money test as ln S < ln K) against basis_log = 1 on
  # exp of the same paths, with half the rows as calibration samples; one
  # asset runs the two-launch date chain, two and five the general one
  # (with the moment products specialised for 2 and 5 assets, ln taken in
  # the product for basis_log = 1)
  from tqf.models import longstaff_schwartz as lsm_lib
  rs = np.random.RandomState(5 + dim)
  n, t = 40000, 9
  z = rs.standard_normal((n, t, dim))
  times = np.linspace(0.1, 0.9, t)
  log_s = np.log(ogbm.sample_paths(times, 0.03, 0.25, 1.0, z))
  dfs = np.exp(-0.03 * times)
  dev = torch.device('cuda')
  lg = torch.as_tensor(log_s, device=dev)
  df = np.concatenate([[1.0], dfs]).reshape(-1, 1, 1)
  for kind, k in ((lsm_lib.PAYOFF_BASKET_PUT, 1.05),
                  (lsm_lib.PAYOFF_MAX_CALL, 0.95)):
    got = lsm_lib._native_lsm(lg, list(range(t)), (kind, np.array([k])), 2,
                              df, n // 2, torch.float64, dev, basis_log=2)
    want = lsm_lib._native_lsm(torch.exp(lg), list(range(t)),
                               (kind, np.array([k])), 2, df, n // 2,
                               torch.float64, dev, basis_log=1)
    np.testing.assert_allclose(_np(got), _np(want), rtol=1e-10)
    # and the oracle on the same rows (regression on the first half)
    def payoff(x, ti, kind=kind, k=k):
      s = np.exp(x[:, ti, :])
      v = (k - s.mean(-1) if kind == lsm_lib.PAYOFF_BASKET_PUT
           else s.max(-1) - k)
      return np.maximum(v, 0)[:, None]
    ref = olsm.least_square_mc(log_s, list(range(t)), payoff,
                               olsm.make_polynomial_basis(2, True),
                               discount_factors=dfs,
                               num_calibration_samples=n // 2)[0]
    np.testing.assert_allclose(float(_np(got)[0]), ref, rtol=2e-6)


@pytest.mark.parametrize('degree', [0, 1, 2, 3, 4])
def test_one_asset_chain_every_degree_matches_oracle(degree):
  # the two-launch one-asset date chain (degree <= 4, full basis) against
  # the oracle's least_square_mc on the same paths, with calibration rows
  # and per-date discount factors
  from tqf.models import longstaff_schwartz as lsm_lib
  rs = np.random.RandomState(20 + degree)
  n, t = 60000, 7
  z = rs.standard_normal((n, t, 1))
  times = np.linspace(0.15, 1.05, t)
  paths = ogbm.sample_paths(times, 0.04, 0.3, 1.0, z)
  dfs = np.exp(-0.04 * times)
  dev = torch.device('cuda')
  got = lsm_lib._native_lsm(
      torch.as_tensor(paths, device=dev), list(range(t)),
      (lsm_lib.PAYOFF_BASKET_PUT, np.array([1.05])), degree,
      np.concatenate([[1.0], dfs]).reshape(-1, 1, 1), n // 3,
      torch.float64, dev)
  want = olsm.least_square_mc(
      paths, list(range(t)),
      lambda x, ti: np.maximum(1.05 - x[:, ti, :1], 0),
      olsm.make_polynomial_basis(degree, True), discount_factors=dfs,
      num_calibration_samples=n // 3)[0]
  np.testing.assert_allclose(float(_np(got)[0]), want, rtol=2e-6)
