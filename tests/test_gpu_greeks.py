"""GPU batched pricing and adjoint Greeks against closed forms, the CPU
oracle with common random numbers, and central finite differences."""
import numpy as np
import pytest
import torch

from oracle import closed_form, gbm as ogbm, heston as oheston, rng as orng

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
  from tqf.engine import batch as batch_lib
  from tqf.engine import greeks


def _gbm(i, s0=100., k=100., t=1.0, sigma=0.25, steps=50, paths=2**16):
  return dict(kind=batch_lib.KIND_GBM, n_steps=steps, n_paths=paths, s0=s0,
              strike=k, maturity=t, rate=0.03, sigma=sigma, v0=0., kappa=0.,
              theta=0., xi=0., rho=0., stream_id=i)


def _heston(i, s0=100., k=100., t=1.0, steps=60, paths=4096, v0=0.04,
            kappa=1.5, theta=0.05, xi=0.5, rho=-0.6):
  return dict(kind=batch_lib.KIND_HESTON, n_steps=steps, n_paths=paths, s0=s0,
              strike=k, maturity=t, rate=0.03, sigma=0., v0=v0, kappa=kappa,
              theta=theta, xi=xi, rho=rho, stream_id=i)


def test_batch_gbm_prices_and_greeks_vs_black_scholes():
  cs = [_gbm(i, s0=80 + 10 * i, k=100., t=0.5 + 0.25 * i,
             sigma=0.15 + 0.05 * i) for i in range(6)]
  plan = batch_lib.BatchPlan(cs, tile=1024)
  plan.launch()
  res = plan.results()
  for c, r in zip(cs, res):
    args = (c['s0'], c['strike'], c['maturity'], c['rate'], c['sigma'])
    for name, fn in (('price', closed_form.bs_price),
                     ('delta', closed_form.bs_delta),
                     ('vega', closed_form.bs_vega),
                     ('rho', closed_form.bs_rho)):
      m, se = r[name]
      assert abs(m - fn(*args)) <= 4 * se + 1e-4 * abs(fn(*args)), (name, m,
                                                                     fn(*args),
                                                                     se)


def _oracle_heston(c, z, bump=None, h=0.0):
  p = dict(c)
  if bump is not None:
    p[bump] = p[bump] + h
  n = c['n_steps']
  t = np.linspace(0, p['maturity'], n + 1)
  keep = np.zeros(n + 1, bool)
  keep[-1] = True
  pr, _ = oheston.ft_euler_calls(np.log(p['s0']), p['v0'], p['kappa'],
                                 p['theta'], p['xi'], p['rho'], p['rate'],
                                 [p['strike']], t, keep, z)
  return pr[-1, 0]


@pytest.mark.parametrize('greeks', ['adjoint', 'tangent'])
def test_batch_heston_price_and_adjoint_vs_oracle_fd(greeks):
  cs = [_heston(0), _heston(1, k=110., xi=0.8, rho=-0.3, steps=45),
        _heston(2, k=90., v0=0.09, kappa=3.0, steps=33)]
  plan = batch_lib.BatchPlan(cs, tile=1024, seed=5, greeks=greeks)
  plan.launch()
  res = plan.results()
  for c, r in zip(cs, res):
    z = orng.normals(c['n_paths'], 2 * c['n_steps'], 5,
                     stream=c['stream_id'], dtype=np.float32)
    z = z.reshape(c['n_paths'], c['n_steps'], 2)
    price = _oracle_heston(c, z)
    assert abs(r['price'][0] - price) <= 2e-4 * price + 1e-4
    # Central differences of the float64 oracle with common random numbers.
    # The steps must be tiny: sqrt(v+) is far from linear for the paths whose
    # variance sits near 0, so h = 1e-6 on theta is already outside the
    # linear range (it converges to the pathwise value only for h <= 1e-7).
    for name, field, h in (('delta', 's0', 1e-5), ('dv0', 'v0', 1e-9),
                           ('rho', 'rate', 1e-9), ('dkappa', 'kappa', 1e-8),
                           ('dtheta', 'theta', 1e-9), ('dxi', 'xi', 1e-8),
                           ('drho', 'rho', 1e-8)):
      fd = (_oracle_heston(c, z, field, h) - _oracle_heston(c, z, field, -h)) / (
          2 * h)
      # Same estimator, same draws: the float32 kernel and the float64
      # oracle differ by rounding, which near v = 0 moves the 1/sqrt(v)
      # factors of a few paths; the bound is 0.2 % or a tenth of the
      # estimator's own standard error (Feller-violating contracts have very
      # heavy-tailed pathwise variance Greeks).
      got, se = r[name]
      scale = max(abs(fd), 1e-2 * abs(price))
      assert abs(got - fd) <= 2e-3 * scale + 0.1 * se, (name, got, fd, se)


def test_batch_tangent_greeks_equal_adjoint_greeks():
  # the forward tangents and the checkpointed reverse sweep compute the
  # same pathwise estimator on the same draws: equal up to float32
  # rounding (prices identical)
  cs = [_heston(0), _heston(1, k=110., xi=0.8, rho=-0.3, steps=45),
        _heston(2, k=90., v0=0.09, kappa=3.0, steps=33),
        _heston(3, steps=121, paths=3000), _gbm(4)]
  adj = batch_lib.BatchPlan(cs, tile=512, seed=3).launch().cpu().numpy()
  tan = batch_lib.BatchPlan(cs, tile=512, seed=3,
                            greeks='tangent').launch().cpu().numpy()
  np.testing.assert_array_equal(tan[:, 0], adj[:, 0])
  means = slice(0, 8)
  scale = np.maximum(np.abs(adj[:, means]), 1e-3 * np.abs(adj[:, :1]))
  assert np.all(np.abs(tan[:, means] - adj[:, means]) <= 2e-4 * scale), (
      np.max(np.abs(tan[:, means] - adj[:, means]) / scale))


def test_batch_contract_longer_than_the_checkpoint_workspace_takes_tangents():
  # a C caller may size the checkpoint workspace for fewer steps than an
  # adjoint-mode Heston contract takes: that contract falls back to the
  # forward tangents (same estimator) and nothing is written past the
  # workspace; the others keep the reverse sweep
  cs = [_heston(0, steps=200, paths=3000), _heston(1, steps=40, paths=3000)]
  short = batch_lib.BatchPlan(cs, tile=512, seed=6)
  assert short.max_steps == 200
  short.max_steps = 48  # checkpoints for 3 segments of 16 steps
  got = short.launch().cpu().numpy()
  tan = batch_lib.BatchPlan(cs, tile=512, seed=6,
                            greeks='tangent').launch().cpu().numpy()
  adj = batch_lib.BatchPlan(cs, tile=512, seed=6).launch().cpu().numpy()
  np.testing.assert_array_equal(got[:, 0], adj[:, 0])
  np.testing.assert_allclose(got[0], tan[0], rtol=1e-6, atol=1e-9)
  np.testing.assert_array_equal(got[1], adj[1])


def test_batch_equals_loop_of_single_contracts():
  cs = [_gbm(0), _heston(1), _gbm(2, sigma=0.4, steps=77, paths=5000),
        _heston(3, steps=121, paths=3000)]
  plan = batch_lib.BatchPlan(cs, tile=512, seed=9)
  out = plan.launch().cpu().numpy()
  for i, c in enumerate(cs):
    one = batch_lib.BatchPlan([c], tile=512, seed=9).launch().cpu().numpy()
    np.testing.assert_allclose(out[i], one[0], rtol=1e-6, atol=1e-9)


def test_batch_graph_replay_matches_eager():
  cs = [_gbm(i, steps=40 + i, paths=3000 + 100 * i) for i in range(3)] + [
      _heston(3, paths=2000)]
  plan = batch_lib.BatchPlan(cs, tile=512, seed=4)
  eager = plan.launch().clone()
  g = plan.capture()
  for _ in range(3):
    plan.out.zero_()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    torch.testing.assert_close(plan.out, eager, rtol=1e-7, atol=1e-12)
    _ = torch.ones(1000, device=plan.dev).sum()  # eager work between replays


def test_basket_adjoint_vs_oracle_pathwise_and_fd():
  d, steps, n = 16, 64, 2**14
  rs = np.random.RandomState(3)
  a = rs.standard_normal((d, d))
  c = a @ a.T + np.eye(d)
  corr = c / np.sqrt(np.outer(np.diag(c), np.diag(c)))
  sig = rs.uniform(0.15, 0.45, d)
  s0 = np.full(d, 100.0)
  res = greeks.basket_call_greeks(s0, sig, corr, 0.02, 1.0, 100.0, steps, n,
                                  seed=3)
  z = orng.normals(n, steps * d, 3, stream=0,
                   dtype=np.float32).reshape(n, steps, d)
  p = ogbm.basket_call_paths(s0, sig, corr, 0.02, 1.0, 100.0, z)
  np.testing.assert_allclose(res['price'][0], p['pay'].mean(), rtol=2e-4)
  np.testing.assert_allclose(res['rho'][0], p['rho'].mean(), rtol=2e-4)
  np.testing.assert_allclose(res['delta'][0], p['delta'].mean(0), rtol=5e-4,
                             atol=1e-6)
  np.testing.assert_allclose(res['vega'][0], p['vega'].mean(0), rtol=5e-4,
                             atol=1e-4)
  # central differences with common random numbers (h = 1e-4 relative)
  for i in (0, 7, 15):
    e = np.zeros(d)
    e[i] = 1e-2
    fd = (ogbm.basket_call_price(s0 + e, sig, corr, .02, 1., 100., z) -
          ogbm.basket_call_price(s0 - e, sig, corr, .02, 1., 100., z)) / 2e-2
    np.testing.assert_allclose(res['delta'][0][i], fd, rtol=2e-3)


def test_autograd_functions_return_adjoint_greeks():
  s0 = torch.tensor(100.0, dtype=torch.float64, requires_grad=True)
  v0 = torch.tensor(0.04, dtype=torch.float64, requires_grad=True)
  r = torch.tensor(0.03, dtype=torch.float64, requires_grad=True)
  k, th, xi, rho = (torch.tensor(v, dtype=torch.float64, requires_grad=True)
                    for v in (1.5, 0.05, 0.5, -0.6))
  cfg = dict(n_steps=60, strike=100.0, maturity=1.0, n_paths=4096, seed=5,
             stream_id=0)
  price = greeks.HestonCall.apply(s0, v0, r, k, th, xi, rho, cfg)
  grads = torch.autograd.grad(price, [s0, v0, r, k, th, xi, rho])
  plan = batch_lib.BatchPlan([_heston(0)], tile=1024, seed=5)
  direct = plan.launch()[0, 1:8].cpu().numpy()
  np.testing.assert_allclose([g.item() for g in grads], direct, rtol=1e-12)
  # basket
  d = 4
  s = torch.full((d,), 100.0, dtype=torch.float64, requires_grad=True)
  sg = torch.tensor([0.2, 0.25, 0.3, 0.35], dtype=torch.float64,
                    requires_grad=True)
  rr = torch.tensor(0.02, dtype=torch.float64, requires_grad=True)
  corr = 0.5 * np.eye(d) + 0.5
  p = greeks.basket_call(s, sg, rr, corr=corr, maturity=1.0, strike=100.0,
                         n_steps=16, n_paths=2**15, seed=1)
  gs, gsig, gr = torch.autograd.grad(p, [s, sg, rr])
  ref = greeks.basket_call_greeks(s.detach().cpu().numpy(),
                                  sg.detach().cpu().numpy(), corr, 0.02, 1.0,
                                  100.0, 16, 2**15, seed=1)
  np.testing.assert_allclose(gs.cpu().numpy(), ref['delta'][0], rtol=1e-12)
  np.testing.assert_allclose(gsig.cpu().numpy(), ref['vega'][0], rtol=1e-12)
  np.testing.assert_allclose(gr.item(), ref['rho'][0], rtol=1e-12)


def test_watch_params_gradients_through_the_generic_loop():
  # euler_sampling.sample with user callables that close over a tensor in
  # watch_params (the reference's forward-gradient route,
  # euler_sampling.py:467-510) is differentiable end to end: the pathwise
  # vega equals central differences with common random numbers and lies
  # within Monte Carlo error of Black-Scholes.
  import tqf
  from tqf.math.random import multivariate_normal as mvn
  r, k, n = 0.05, 100.0, 100000

  def price(sig):
    paths = tqf.models.euler_sampling.sample(
        1, lambda t, x: r * x, lambda t, x: sig * x.unsqueeze(-1), [1.0],
        num_time_steps=50, num_samples=n, initial_state=[100.0],
        random_type=mvn.RandomType.PSEUDO, seed=3, watch_params=[sig],
        dtype=torch.float64)
    pay = torch.relu(paths[:, -1, 0] - k) * np.exp(-r)
    return pay.mean(), pay

  sig = torch.tensor(0.2, dtype=torch.float64, device='cuda',
                     requires_grad=True)
  p, pay = price(sig)
  p.backward()
  vega = float(sig.grad)
  h = 1e-6
  with torch.no_grad():
    up = float(price(torch.tensor(0.2 + h, dtype=torch.float64,
                                  device='cuda'))[0])
    dn = float(price(torch.tensor(0.2 - h, dtype=torch.float64,
                                  device='cuda'))[0])
  fd = (up - dn) / (2 * h)
  assert abs(vega - fd) <= 1e-5 * abs(fd) + 1e-6, (vega, fd)
  bs = closed_form.bs_vega(100.0, k, 1.0, r, 0.2)
  assert abs(vega - bs) <= 0.03 * bs, (vega, bs)


def test_notebook_greeks_through_generic_ito_process_and_fwd_gradient():
  # Monte_Carlo_Euler_Scheme.ipynb (cells 10 and 22-27): log-GBM through
  # GenericItoProcess with callables that close over sigma, watch_params=
  # [sigma], PSEUDO_ANTITHETIC draws, time_step 0.01, 200,000 paths,
  # expiries [0.5, 1.0] x strikes [600, 650, 680]; delta and vega by
  # tqf.math.fwd_gradient against Black-Scholes.  The notebook's own
  # maximum relative errors (delta 1.71e-3, vega 7.78e-2, :917-920) set the
  # scale of the bound: a different random stream, the same estimator.
  print('notebook Greeks via fwd_gradient', flush=True)
  import tqf
  from conftest import load_golden
  from tqf.math.random import multivariate_normal as mvn
  g = load_golden('notebook_euler.json')
  dev = 'cuda'
  f64 = dict(dtype=torch.float64, device=dev)
  rate = 0.03
  expiries = torch.tensor([0.5, 1.0], **f64)
  strikes = torch.tensor([600.0, 650.0, 680.0], **f64)
  spot = torch.tensor(700.0, **f64)
  sigma = torch.tensor(0.1, **f64)

  def price_eu_options(strikes, spot, sigma):
    def drift_fn(t, x):
      del t, x
      return rate - 0.5 * sigma**2

    def vol_fn(t, x):
      del t, x
      return torch.reshape(sigma, [1, 1])
    process = tqf.models.GenericItoProcess(dim=1, drift_fn=drift_fn,
                                           volatility_fn=vol_fn,
                                           dtype=torch.float64)
    paths = process.sample_paths(
        [0.5, 1.0], num_samples=200000,
        initial_state=torch.log(torch.mean(spot)), watch_params=[sigma],
        random_type=mvn.RandomType.PSEUDO_ANTITHETIC, seed=42, time_step=0.01)
    return (torch.exp(-rate * expiries).unsqueeze(-1) *
            torch.mean(torch.relu(torch.exp(paths) - strikes), 0))

  prices = price_eu_options(strikes, spot, sigma)
  assert prices.shape == (2, 3)
  delta = tqf.math.fwd_gradient(
      lambda s: price_eu_options(strikes, s, sigma), spot)
  vega = tqf.math.fwd_gradient(
      lambda v: price_eu_options(strikes, spot, v), sigma)
  t = np.array([[0.5], [1.0]])
  k = np.array([[600.0, 650.0, 680.0]])
  bs = closed_form.bs_price(700.0, k, t, rate, 0.1)
  bs_delta = closed_form.bs_delta(700.0, k, t, rate, 0.1)
  bs_vega = closed_form.bs_vega(700.0, k, t, rate, 0.1)
  np.testing.assert_allclose(prices.detach().cpu().numpy(), bs, rtol=2e-3)
  err_d = np.max(np.abs(delta.cpu().numpy() / bs_delta - 1))
  err_v = np.max(np.abs(vega.cpu().numpy() / bs_vega - 1))
  print('max relative error: delta %.2e (notebook %.2e), vega %.2e '
        '(notebook %.2e)' % (err_d, g['max_rel_err_delta'], err_v,
                             g['max_rel_err_vega']), flush=True)
  assert err_d <= 3 * g['max_rel_err_delta']
  assert err_v <= 3 * g['max_rel_err_vega']
  # the reverse-mode route gives the same sums over the output grid
  d_sum = tqf.math.gradients(lambda s: price_eu_options(strikes, s, sigma),
                             spot)
  np.testing.assert_allclose(float(d_sum), float(delta.sum()), rtol=1e-10)


def test_generic_ito_process_routes_model_callables_to_the_fused_kernel():
  import tqf
  gbm = tqf.models.GeometricBrownianMotion(0.05, 0.2, dtype=torch.float64)
  proc = tqf.models.GenericItoProcess(1, gbm.drift_fn(), gbm.volatility_fn(),
                                      dtype=torch.float64)
  z = np.random.RandomState(4).standard_normal((1000, 20, 1))
  a = proc.sample_paths([0.5, 1.0], initial_state=[100.0], num_time_steps=20,
                        normal_draws=z)
  b = tqf.models.euler_sampling.sample(
      1, lambda t, x: 0.05 * x, lambda t, x: 0.2 * x.unsqueeze(-1),
      [0.5, 1.0], initial_state=[100.0], num_time_steps=20, normal_draws=z,
      dtype=torch.float64)
  np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=1e-12)
  with pytest.raises(NotImplementedError):
    proc.fd_solver_backward(1.0, 0.0, None, None)
  with pytest.raises(ValueError):
    tqf.models.GenericItoProcess(0, gbm.drift_fn(), gbm.volatility_fn())


def test_gbm_sample_paths_gradients_match_the_per_path_derivatives():
  # GeometricBrownianMotion.sample_paths is differentiable in S0, mu and
  # sigma (the reference tapes it for Greeks): against the per-path
  # derivatives of the exact scheme on the same draws
  import tqf
  rs = np.random.RandomState(6)
  n, times = 4000, np.array([0.25, 0.6, 1.0])
  z = rs.standard_normal((n, 3, 1))
  w = rs.standard_normal((n, 3, 1))
  f64 = dict(dtype=torch.float64, device='cuda')
  mu = torch.tensor(0.04, requires_grad=True, **f64)
  sig = torch.tensor(0.3, requires_grad=True, **f64)
  s0 = torch.tensor(100.0, requires_grad=True, **f64)
  gbm = tqf.models.GeometricBrownianMotion(mu, sig, dtype=torch.float64)
  p = gbm.sample_paths(times, initial_state=s0, normal_draws=z)
  got = torch.autograd.grad((p * torch.as_tensor(w, device='cuda')).sum(),
                            [s0, mu, sig])
  s = p.detach().cpu().numpy()
  t = times.reshape(1, -1, 1)
  wt = np.cumsum(np.sqrt(np.diff(np.concatenate([[0.0], times]))).reshape(
      1, -1, 1) * z, axis=1)
  want = [np.sum(w * s / 100.0), np.sum(w * s * t),
          np.sum(w * s * (wt - 0.3 * t))]
  np.testing.assert_allclose([float(g) for g in got], want, rtol=1e-10)
  # a call's pathwise vega from the taped sampler against Black-Scholes
  sig2 = torch.tensor(0.2, requires_grad=True, **f64)
  g2 = tqf.models.GeometricBrownianMotion(0.03, sig2, dtype=torch.float64)
  st = g2.sample_paths([1.0], initial_state=100.0, num_samples=2**20,
                       seed=4)[:, -1, 0]
  price = torch.exp(torch.tensor(-0.03, **f64)) * torch.relu(st - 100).mean()
  vega = float(torch.autograd.grad(price, sig2)[0])
  bs = closed_form.bs_vega(100.0, 100.0, 1.0, 0.03, 0.2)
  assert abs(vega - bs) < 0.01 * bs, (vega, bs)


@pytest.mark.parametrize('greeks', ['tangent', 'adjoint'])
def test_batch_table_sample_price_and_greeks_vs_oracle(greeks):
  # 32 contracts of the heterogeneous 4096-contract table (every 128th:
  # GBM and Heston, 50-500 steps, maturities 0.1-3), at most 2048 paths
  # each: price and every Greek against the float64 oracle on the same
  # Philox draws (Heston: forward tangents of the step loop; GBM: the
  # closed-form pathwise terms)
  print('batch table sample vs oracle', greeks, flush=True)
  table = batch_lib.random_contracts(4096)
  cs = []
  for i in range(0, 4096, 128):
    c = dict(table[i])
    c['n_paths'] = min(c['n_paths'], 2048)
    c['stream_id'] = i
    cs.append(c)
  plan = batch_lib.BatchPlan(cs, tile=1024, seed=5, greeks=greeks)
  plan.launch()
  res = plan.results()
  for c, r in zip(cs, res):
    n, m = c['n_steps'], c['n_paths']
    if c['kind'] == batch_lib.KIND_GBM:
      z = orng.normals(m, n, 5, stream=c['stream_id'],
                       dtype=np.float32).astype(np.float64)
      dt = c['maturity'] / n
      w = np.sqrt(dt) * z.sum(1)
      s = c['s0'] * np.exp((c['rate'] - 0.5 * c['sigma']**2) * c['maturity'] +
                           c['sigma'] * w)
      disc = np.exp(-c['rate'] * c['maturity'])
      itm = s > c['strike']
      pay = disc * np.where(itm, s - c['strike'], 0)
      sb = disc * np.where(itm, s, 0)
      g = np.stack([pay, sb / c['s0'], sb * (w - c['sigma'] * c['maturity']),
                    -c['maturity'] * pay + sb * c['maturity']])
      want, want_se = g.mean(1), g.std(1, ddof=1) / np.sqrt(m)
    else:
      z = orng.normals(m, 2 * n, 5, stream=c['stream_id'],
                       dtype=np.float32).astype(np.float64).reshape(m, n, 2)
      want, want_se = oheston.ft_euler_call_greeks(
          np.log(c['s0']), c['v0'], c['kappa'], c['theta'], c['xi'], c['rho'],
          c['rate'], c['strike'], np.linspace(0.0, c['maturity'], n + 1), z)
      # kernel order: price, delta, d/dv0, rho_r, kappa, theta, xi, rho
    names = batch_lib.GREEK_NAMES[c['kind']]
    if (c['kind'] == batch_lib.KIND_HESTON and
        2 * c['kappa'] * c['theta'] < c['xi']**2):
      # Feller condition violated: the variance reaches 0 and the pathwise
      # variance-parameter Greeks carry 1/sqrt(v) factors with unbounded
      # tails (float32 and float64 then differ by more than their standard
      # errors); price, delta and rate rho stay well conditioned
      names = names[:2] + names[3:4]
    price = want[0]
    for name in names:
      q = batch_lib.GREEK_NAMES[c['kind']].index(name)
      got, se = r[name]
      scale = max(abs(want[q]), 1e-2 * abs(price), 1e-6)
      assert abs(got - want[q]) <= 2e-3 * scale + 0.1 * want_se[q], (
          c['kind'], name, got, want[q], want_se[q])
