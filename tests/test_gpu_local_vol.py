"""Local volatility on the GPU: the edge-mode table in the fused kernel
against the oracle, Dupire surfaces from market data reproducing their
implied-volatility smile (local_volatility_model_test.py:102-378), and the
fused precomputed surface against the per-step formula on the same draws.
"""
import numpy as np
import pytest
import torch
from scipy import optimize

from conftest import load_golden
from oracle import closed_form
from oracle import euler as oeuler
from oracle import local_vol as olv

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
  import tqf
  from tqf.experimental import volatility_surface as vsl
  from tqf.math.random import multivariate_normal as mvn

G = load_golden('interpolation_dupire.json')['dupire']


def _np(t):
  return t.detach().cpu().numpy()


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12),
                                       (torch.float32, 2e-5)])
def test_edge_mode_table_parity(dtype, tol):
  npd = np.float64 if dtype == torch.float64 else np.float32
  times = np.array([0.0, 0.4, 1.0])
  x = np.linspace(np.log(80.0), np.log(125.0), 7)
  rs = np.random.RandomState(2)
  vols = 0.2 + 0.05 * rs.uniform(size=(3, 7))
  outside = np.stack([0.35 + 0.01 * np.arange(3),
                      0.12 + 0.01 * np.arange(3)], 1)
  surf = tqf.models.LocalVolSurface(times, x[0], x[1] - x[0], vols,
                                    outside=outside)
  model = tqf.models.LocalVolatilityModel(1, risk_free_rate=0.02,
                                          local_volatility_fn=surf,
                                          dtype=dtype)
  n, steps = 4000, 30
  z = rs.standard_normal((n, steps, 1)).astype(npd)
  got = _np(model.sample_paths([0.5, 1.0], initial_state=[100.0],
                               num_time_steps=steps, normal_draws=z))
  fns = olv.lv_fns(times, x[0], x[1] - x[0], vols, 0.02, 0.0, outside)
  want = oeuler.sample(1, fns[0], fns[1], np.asarray([0.5, 1.0], npd),
                       num_samples=n,
                       initial_state=np.asarray([np.log(100.)], npd),
                       normal_draws=z, num_time_steps=steps, dtype=npd)
  # paths cross both ends of the table
  assert (want < x[0]).any() and (want > x[-1]).any()
  np.testing.assert_allclose(got, np.exp(want), rtol=tol, atol=0)


def _implied_vol(price, s0, k, t, df):
  f = s0 / df

  def bs(sig):
    return closed_form.bs_price(f * df, k, t, -np.log(df) / t, sig) - price
  return optimize.brentq(bs, 1e-4, 3.0, xtol=1e-10)


def _market(dim, scale=1.0):
  strikes = np.arange(0.9, 1.3, 0.04) * scale
  iv = np.arange(0.21, 0.31, 0.01)
  expiries = vsl.dates_from_year_month_day(dim * [[2021, 2022]],
                                           dim * [[1, 1]], dim * [[1, 1]])
  et = np.array(G['expiry_days']) / 365.0
  return (expiries, np.broadcast_to(strikes, (dim, 2, 10)).copy(),
          np.broadcast_to(iv, (dim, 2, 10)).copy(), et)


def _callable_df(lower, upper):
  def df(t):
    t = torch.as_tensor(t, dtype=torch.float64)
    v = torch.where(t < 0.5, torch.exp(-lower * t),
                    torch.exp(-lower * 0.5 - upper * (t - 0.5)))
    return v.unsqueeze(-1)
  return df


def _const_df(r):
  rt = torch.tensor([r], dtype=torch.float64)
  return lambda t: torch.exp(-rt * torch.as_tensor(t, dtype=torch.float64))


def _check_smile(paths, expiries, strikes, iv, spot, df_fn, start=0):
  for d in range(paths.shape[-1]):
    for i, t in enumerate(expiries):
      df = float(_np(torch.as_tensor(df_fn(torch.tensor(t)))).reshape(-1)[0])
      st = paths[:, i, d]
      for j in range(start, strikes.shape[-1]):
        k = strikes[d, i, j]
        val = max(np.maximum(st - k, 0).mean(), spot - k)
        sim = _implied_vol(df * val, spot, k, t, df)
        np.testing.assert_allclose(sim, iv[d, i, j], atol=0.05, rtol=0.005)


@pytest.mark.parametrize('case', [
    ('1d', 1, 0.0, None, 20, 0),
    ('2d', 2, 0.0, None, 20, 0),
    ('1d_nonzero_rate', 1, 0.05, None, 40, 1),
    ('1d_callable_rate', 1, None, (0.01, 0.02), 40, 1),
])
def test_from_market_data_reproduces_smile(case):
  _, dim, r, rates, steps, start = case
  expiries, strikes, iv, et = _market(dim)
  df = _const_df(r) if rates is None else _callable_df(*rates)
  lv = tqf.models.LocalVolatilityModel.from_market_data(
      dim, [(2020, 1, 1)], expiries, strikes, iv, [1.0] * dim, df,
      dividend_yield=[0.0], dtype=torch.float64)
  paths = _np(lv.sample_paths(
      et, num_samples=10000, initial_state=[1.0] * dim,
      num_time_steps=steps,
      random_type=mvn.RandomType.STATELESS_ANTITHETIC, seed=[1, 2]))
  assert paths.shape == (10000, 2, dim)
  _check_smile(paths, et, strikes, iv, 1.0, df, start)


def _precomputed_model(r, num_steps=20, nodes=2049):
  expiries, strikes, iv, et = _market(1)
  tg = np.sort(np.concatenate([np.arange(0.0, 3.0, 3.0 / num_steps), et]))
  sg = np.arange(0.9, 1.3, 0.04)
  lv = tqf.models.LocalVolatilityModel.from_market_data(
      1, [(2020, 1, 1)], expiries, strikes, iv, [1.0], _const_df(r),
      dividend_yield=[0.0], times_grid=tg, spot_grid=sg, precompute_iv=True,
      dtype=torch.float64, num_spot_nodes=nodes)
  return lv, et, strikes, iv, tg


def test_precomputed_fused_matches_per_step_formula():
  # zero rates: the surface's local volatilities stay in [0.16, 0.38]
  # (with rates, the reference's formula gives local volatilities above 10
  # near the spot-grid ends, where single paths explode in both versions)
  lv, et, _, _, tg = _precomputed_model(0.0)
  assert lv._spec() is not None  # the fused kernel takes it
  steps = tg.shape[0] - 1
  rs = np.random.RandomState(5)
  z = rs.standard_normal((20000, steps, 1))
  fused = _np(lv.sample_paths(et, initial_state=[1.0], normal_draws=z))
  # the same model through the generic loop: the precomputed function
  # itself, evaluated per step
  fn = lv.local_volatility_fn()
  slow = tqf.models.LocalVolatilityModel(
      1, risk_free_rate=0.0, local_volatility_fn=lambda t, s: fn(t, s),
      times_grid=tg, spot_grid=[0.9, 1.26], precompute_iv=True,
      dtype=torch.float64)
  assert slow._spec() is None
  exact = _np(slow.sample_paths(et, initial_state=[1.0], normal_draws=z))
  # linear interpolation in ln S between 2049 nodes (3e-8 at most on the
  # host restatement of both)
  assert np.isfinite(fused).all() and np.isfinite(exact).all()
  np.testing.assert_allclose(fused, exact, rtol=1e-6)


@pytest.mark.parametrize('r,start', [(0.0, 0), (0.05, 1)])
def test_precomputed_reproduces_smile(r, start):
  lv, et, strikes, iv, _ = _precomputed_model(r)
  paths = _np(lv.sample_paths(
      et, num_samples=20000, initial_state=[1.0],
      random_type=mvn.RandomType.STATELESS_ANTITHETIC, seed=[1, 2]))
  _check_smile(paths, et, strikes, iv, 1.0, _const_df(r), start)
