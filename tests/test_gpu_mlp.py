"""GPU MLP kernels (bf16 MFMA forward/backward, Adam) against the numpy
oracle, and the learned Longstaff-Schwartz against the polynomial judge."""
import numpy as np
import pytest
import torch

from oracle import gbm as ogbm
from oracle import lsm as olsm
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
  from tqf.engine import mlp_lsm
  from tqf.models import longstaff_schwartz as lsm_lib


def _np(t):
  return t.detach().cpu().numpy()


def _data(n, seed=0):
  rs = np.random.RandomState(seed)
  x = np.zeros((n, 8), np.float32)
  x[:, :6] = rs.standard_normal((n, 6))
  t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]).astype(np.float32)
  w = (rs.uniform(size=n) > 0.25).astype(np.float32)
  return x, t, w


def _params(seed=1):
  rs = np.random.RandomState(seed)
  return (omlp.init_params(seed) +
          0.05 * rs.standard_normal(omlp.N_PARAMS)).astype(np.float32)


def test_mlp_forward_matches_oracle():
  n = 1000 + 7  # a ragged last group of 16
  x, _, _ = _data(n)
  p = _params()
  m = mlp_lsm.MlpModel(params=p)
  got = m.forward(torch.as_tensor(x, device=m.dev)).cpu().numpy()
  want_bf, _ = omlp.forward(p.astype(np.float64), x.astype(np.float64),
                            bf=True)
  want, _ = omlp.forward(p.astype(np.float64), x.astype(np.float64))
  scale = np.max(np.abs(want))
  # bf16 operands as the kernel: float32 accumulation order, and rarely a
  # hidden activation that rounds to the neighbouring bf16 value
  err = np.abs(got - want_bf)
  assert np.max(err) <= 1e-4 * scale
  assert np.quantile(err, 0.99) <= 1e-5 * scale
  # against the exact float64 network: bf16 rounding of the 64 x 64 layer
  assert np.max(np.abs(got - want)) <= 2e-2 * scale


def test_mlp_gradient_matches_oracle():
  n = 4096 + 5
  x, t, w = _data(n, 2)
  p = _params(3)
  m = mlp_lsm.MlpModel(params=p, n_blocks=37)
  dev = m.dev
  g, loss = m.grad(torch.as_tensor(x, device=dev),
                   torch.as_tensor(t, device=dev),
                   torch.as_tensor(w, device=dev), inv_norm=1.0 / n)
  g = g.cpu().numpy()
  lo, go = omlp.loss_and_grad(p.astype(np.float64), x.astype(np.float64),
                              t.astype(np.float64), w.astype(np.float64),
                              1.0 / n, bf=True)
  np.testing.assert_allclose(loss.item(), lo, rtol=1e-4)
  scale = np.max(np.abs(go))
  assert np.max(np.abs(g - go)) <= 2e-4 * scale, np.argmax(np.abs(g - go))
  _, gx = omlp.loss_and_grad(p.astype(np.float64), x.astype(np.float64),
                             t.astype(np.float64), w.astype(np.float64),
                             1.0 / n)
  assert np.max(np.abs(g - gx)) <= 3e-2 * np.max(np.abs(gx))
  # a gradient on a row window [first, first + n)
  g2, _ = m.grad(torch.as_tensor(x, device=dev), torch.as_tensor(t, device=dev),
                 torch.as_tensor(w, device=dev), first=100, n=1000,
                 inv_norm=1e-3)
  _, go2 = omlp.loss_and_grad(p.astype(np.float64), x[100:1100].astype(
      np.float64), t[100:1100].astype(np.float64),
                              w[100:1100].astype(np.float64), 1e-3, bf=True)
  assert np.max(np.abs(g2.cpu().numpy() - go2)) <= 2e-4 * np.max(np.abs(go2))


def test_mlp_gradient_is_deterministic():
  # the block reduction has a fixed order: repeated calls agree bit for bit
  n = 70000
  x, t, w = _data(n, 4)
  m = mlp_lsm.MlpModel(params=_params(5))
  dev = m.dev
  args = [torch.as_tensor(a, device=dev) for a in (x, t, w)]
  g1, l1 = m.grad(*args, inv_norm=1.0 / n)
  g1, l1 = g1.clone(), l1.clone()
  for _ in range(3):
    g2, l2 = m.grad(*args, inv_norm=1.0 / n)
    assert torch.equal(g1, g2) and torch.equal(l1, l2)


def test_mlp_adam_steps_match_oracle_and_training_converges():
  n = 8192
  x, t, w = _data(n, 4)
  p = _params(5)
  m = mlp_lsm.MlpModel(params=p)
  dev = m.dev
  xt, tt, wt = (torch.as_tensor(a, device=dev) for a in (x, t, w))
  po = p.astype(np.float64)
  mo = np.zeros_like(po)
  vo = np.zeros_like(po)
  for k in range(1, 4):
    m.train_step(xt, tt, wt, 0, n, 1.0 / n, 1e-3)
    _, g = omlp.loss_and_grad(po, x.astype(np.float64), t.astype(np.float64),
                              w.astype(np.float64), 1.0 / n, bf=True)
    po, mo, vo = omlp.adam(po, mo, vo, g, k)
  np.testing.assert_allclose(m.params.cpu().numpy(), po, rtol=0, atol=3e-5)
  first = None
  for k in range(300):
    m.train_step(xt, tt, wt, 0, n, 1.0 / n, 3e-3)
    if k == 0:
      first = m.loss.item()
  assert m.loss.item() < 0.2 * first


@pytest.mark.parametrize('dim,basis_log,kind', [
    (1, 0, 'put'), (2, 1, 'max_call'), (3, 0, 'max_call'), (4, 1, 'put'),
    (5, 0, 'max_call'), (5, 1, 'max_call')])
def test_lsm_date_inputs_match_numpy(dim, basis_log, kind):
  # tqf_lsm_date_inputs on its own: 300,001 paths (1,172 blocks of 256, not
  # a multiple, and enough block sums for every thread of the one-workgroup
  # mean kernel to loop)
  print('lsm date inputs', dim, basis_log, kind, flush=True)
  import ctypes
  from tqf import _native
  n, t_idx, ratio, strike = 300001, 1, 0.97, 100.0
  rs = np.random.RandomState(dim + 10 * basis_log)
  paths = 100 * np.exp(0.25 * rs.standard_normal((n, 3, dim)))
  cash = rs.uniform(0, 5, n) * (rs.uniform(size=n) > 0.5)
  vals = rs.uniform(0, 3, n)
  pt = torch.as_tensor(paths, device='cuda').contiguous()
  payoff = (lsm_lib.PAYOFF_BASKET_PUT if kind == 'put' else
            lsm_lib.PAYOFF_MAX_CALL)
  l, keep = mlp_lsm._lsm_struct(pt, payoff, strike, 2)
  l.basis_log = basis_log
  f64 = dict(dtype=torch.float64, device='cuda')
  ct, vt = torch.as_tensor(cash, **f64), torch.as_tensor(vals, **f64)
  exv = torch.empty(n, **f64)
  y = torch.empty(n, **f64)
  mom = torch.full((_native.lib().tqf_lsm_moments_size(n) // 8,), np.nan,
                   **f64)
  x = torch.full((n, 8), np.nan, dtype=torch.float32, device='cuda')
  tgt = torch.empty(n, dtype=torch.float32, device='cuda')
  w = torch.empty(n, dtype=torch.float32, device='cuda')
  _native.call('tqf_lsm_date_inputs', ctypes.byref(l), ctypes.c_int32(t_idx),
               ctypes.c_double(ratio), *[ctypes.c_void_p(a.data_ptr()) for a
                                         in (ct, vt, exv, y, mom, x, tgt, w)],
               _native.stream_ptr())
  s = paths[:, t_idx, :]
  b = np.log(s) if basis_log else s
  mean = b.mean(0)
  var = (b * b).mean(0) - mean**2
  m = _np(mom)
  np.testing.assert_allclose(m[:dim], mean, rtol=1e-12)
  np.testing.assert_allclose(m[8:8 + dim], 1 / np.sqrt(var), rtol=1e-9)
  want_exv = (np.maximum(strike - s.mean(-1), 0) if kind == 'put' else
              np.maximum(s.max(-1) - strike, 0))
  np.testing.assert_allclose(_np(exv), want_exv, rtol=1e-14, atol=1e-13)
  np.testing.assert_array_equal(_np(y), ratio * (cash + vals))
  xs = _np(x)
  want_x = ((b - m[:dim]) * m[8:8 + dim]).astype(np.float32)
  np.testing.assert_array_max_ulp(xs[:, :dim], want_x, maxulp=1)
  assert np.all(xs[:, dim:5] == 0) and np.all(xs[:, 6:] == 0)
  np.testing.assert_array_max_ulp(xs[:, 5],
                                  (_np(exv) / strike).astype(np.float32), 0)
  np.testing.assert_array_equal(_np(tgt),
                                (_np(y) / strike).astype(np.float32))
  np.testing.assert_array_equal(_np(w), (_np(exv) > 0).astype(np.float32))


def _check_learned_lsm(res, df0, n_cal):
  """The paired gain statistic, recomputed on the host from the per-path
  values of both policies, and the acceptance rule applied to it."""
  diff = df0 * (_np(res['mlp_path_values']) - _np(res['poly_path_values']))
  diff = diff[n_cal:]
  np.testing.assert_allclose(res['gain'], diff.mean(), rtol=1e-9, atol=1e-12)
  np.testing.assert_allclose(res['gain_se'],
                             diff.std(ddof=1) / np.sqrt(diff.size), rtol=1e-6)
  np.testing.assert_allclose(res['gain'], res['mlp_price'] - res['poly_price'],
                             rtol=1e-9, atol=1e-9)
  # both prices are means over the same pricing paths: the paired error is
  # far below the error of two independent prices
  assert 0 < res['gain_se'] < res['gain_se_unpaired']
  assert res['accepted'] == (res['gain'] >= -2 * res['gain_se'])
  assert res['price'] == (res['mlp_price'] if res['accepted'] else
                          res['poly_price'])
  assert all(np.isfinite(res['losses']))


def test_mlp_lsm_prices_against_polynomial_judge():
  # 3-asset Bermudan max-call, small sizes
  paths, df = mlp_lsm.bermudan_max_call_paths(2**16, dim=3, n_dates=5,
                                              steps_per_date=4, maturity=1.0)
  res = mlp_lsm.mlp_lsm_price(paths, [1, 2, 3, 4, 5], 100.0, df, 2**15,
                              steps_per_date=60, batch=8192, lr=3e-3,
                              return_path_values=True)
  # the polynomial judge matches the oracle (standardised basis) on the
  # same paths
  p = paths.cpu().numpy()
  want = olsm.least_square_mc(p, [1, 2, 3, 4, 5],
                              olsm.make_basket_call_max_payoff([100.0]),
                              olsm.make_polynomial_basis(2, True),
                              discount_factors=df,
                              num_calibration_samples=2**15)
  np.testing.assert_allclose(res['poly_price'], want[0], rtol=2e-6)
  assert res['mlp_se'] > 0 and res['poly_se'] > 0
  _check_learned_lsm(res, df[0], 2**15)
  # both are lower-bound prices of the same option on the same pricing
  # paths: a reasonable learned policy is within a few percent of the judge
  assert abs(res['gain']) < 0.02 * res['poly_price']


def test_mlp_lsm_five_assets_gain_over_the_polynomial_judge():
  # the learned-LSM workload's shape (5 assets, 9 dates of 20 steps, the
  # 6 -> 64 -> 64 -> 1 network on 6 inputs, 200 Adam steps per date) at
  # 2^16 calibration + 2^16 pricing paths; reports what the network gains
  print('learned LSM, 5 assets', flush=True)
  paths, df = mlp_lsm.bermudan_max_call_paths(2**17)
  res = mlp_lsm.mlp_lsm_price(paths, list(range(1, 10)), 100.0, df, 2**16,
                              batch=16384, lr=1e-3, return_path_values=True)
  print('5-asset learned LSM: mlp %.4f +- %.4f, poly %.4f +- %.4f, '
        'gain %+.4f +- %.4f (paired), accepted %s' % (
            res['mlp_price'], res['mlp_se'], res['poly_price'],
            res['poly_se'], res['gain'], res['gain_se'], res['accepted']),
        flush=True)
  p = paths.cpu().numpy()
  want = olsm.least_square_mc(p, list(range(1, 10)),
                              olsm.make_basket_call_max_payoff([100.0]),
                              olsm.make_polynomial_basis(2, True),
                              discount_factors=df,
                              num_calibration_samples=2**16)
  np.testing.assert_allclose(res['poly_price'], want[0], rtol=2e-6)
  _check_learned_lsm(res, df[0], 2**16)
  assert abs(res['gain']) < 0.03 * res['poly_price']


def test_mlp_date_training_graph_replays_match_eager():
  # one date's training captured as a hipGraph (Adam reset by fill kernels,
  # then the steps) and replayed three times, with eager work on the
  # inputs and a synchronisation between the replays, against the same
  # steps launched one by one: parameters, Adam state and loss bit for bit
  # after every date
  from tqf.engine.mlp_lsm import MlpModel
  rs = np.random.RandomState(3)
  n = 50000
  x = np.zeros((n, 8), np.float32)
  x[:, :6] = rs.standard_normal((n, 6))
  t = (np.tanh(x[:, 0]) + 0.3 * x[:, 1] * x[:, 2]).astype(np.float32)
  w = (rs.uniform(size=n) > 0.3).astype(np.float32)
  xt, tt, wt = (torch.as_tensor(a, device='cuda') for a in (x, t, w))
  a = MlpModel(seed=1)
  b = MlpModel(seed=1)
  for date in range(3):
    a.train_date(xt, tt, wt, 40000, 8192, 25, 1e-3, graph=True)
    b.train_date(xt, tt, wt, 40000, 8192, 25, 1e-3, graph=False)
    torch.cuda.synchronize()
    for u, v in ((a.params, b.params), (a.m, b.m), (a.v, b.v),
                 (a.loss, b.loss)):
      assert torch.equal(u, v), date
    tt.mul_(0.9).add_(0.05)
    torch.cuda.synchronize()
  assert len(a._graphs) == 1
