"""Differentiable fused sampling: euler_sampling.sample with gradients taken
by the forward-tangent kernel (tqf_euler_tangent), against central
differences of the CPU oracle on the same draws, against autograd through
the per-step loop (the reference's tape route, euler_sampling.py:405-464),
and through tqf.math.fwd_gradient (the JVP)."""
import numpy as np
import pytest
import torch

from oracle import euler as oeuler
from oracle import heston as oheston

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
  import tqf
  from tqf.engine import euler as engine


def _np(t):
  return t.detach().cpu().numpy()


TIMES = [0.3, 0.7, 1.0]
STEPS = 30


def _oracle_fns(kind, p):
  if kind == 'gbm':
    return oeuler.gbm_fns(p[0], p[1])
  if kind == 'const':
    return (lambda t, x: np.full_like(x, p[0]),
            lambda t, x: np.full(x.shape + (1,), p[1]))
  if kind == 'heston':
    return oeuler.heston_ft_fns(p[1], p[2], p[3], p[4], p[0])

  def drift(t, x):
    # HestonModel's own coefficients; the engine adds the rate to the
    # log-spot drift (zero in HestonModel itself)
    return oheston.drift_vol(x, p[1], p[2], p[3], p[4])[0] + np.array(
        [p[0], 0.0])

  def vol(t, x):
    return oheston.drift_vol(x, p[1], p[2], p[3], p[4])[1]
  return drift, vol


CASES = {
    # kind, parameters, x0, engine variant
    'gbm': ('gbm', [0.05, 0.25], [100.0], None),
    'const': ('const', [0.02, 0.3], [np.log(100.0)], None),
    'heston': ('heston', [0.03, 2.0, 0.05, 0.3, -0.6], [np.log(100.0), 0.04],
               0.0),
    'heston_own': ('heston_own', [0.01, 1.5, 0.05, 0.25, -0.5],
                   [np.log(100.0), 0.06], 1.0),
    # variance far from 0 on every path: autograd through sqrt(v+) in the
    # step loop is finite
    'heston_smooth': ('heston', [0.03, 2.0, 0.05, 0.1, -0.6],
                      [np.log(100.0), 0.04], 0.0),
}


def _spec(kind, params, variant):
  if kind.startswith('heston'):
    return ('heston',) + tuple(params) + (variant,)
  return (kind,) + tuple(params)


def _fused(kind, params, x0, variant, z, dtype=torch.float64):
  spec = _spec(kind, params, variant)
  d = engine.model_dim(spec)
  f = engine.tag(lambda t, x: None, spec)
  g = engine.tag(lambda t, x: None, spec)
  return tqf.models.euler_sampling.sample(
      d, f, g, TIMES, num_time_steps=STEPS, initial_state=x0, normal_draws=z,
      dtype=dtype)


def _oracle_paths(kind, params, x0, z):
  d = len(x0)
  fns = _oracle_fns(kind, params)
  return oeuler.sample(d, fns[0], fns[1], np.asarray(TIMES), num_samples=z.shape[0],
                       initial_state=np.asarray(x0), normal_draws=z,
                       num_time_steps=STEPS)


@pytest.mark.parametrize('name', ['gbm', 'const', 'heston', 'heston_own'])
def test_tangent_vjp_matches_oracle_central_differences(name):
  print('tangent vjp vs oracle fd', name, flush=True)
  kind, p0, x00, variant = CASES[name]
  d = len(x00)
  rs = np.random.RandomState(3)
  n = 2000
  z = rs.standard_normal((n, STEPS + 2, d))
  w = rs.standard_normal((n, len(TIMES), d))
  dev = 'cuda'
  x0 = torch.tensor(x00, dtype=torch.float64, device=dev, requires_grad=True)
  params = [torch.tensor(v, dtype=torch.float64, device=dev,
                         requires_grad=True) for v in p0]
  paths = _fused(kind, params, x0, variant, z)
  assert paths.requires_grad and paths.shape == (n, len(TIMES), d)
  # the forward is the fused kernel's own output
  np.testing.assert_allclose(_np(paths), _fused(kind, p0, x00, variant, z)
                             .cpu().numpy(), rtol=0, atol=0)
  loss = (paths * torch.as_tensor(w, device=dev)).sum()
  got = torch.autograd.grad(loss, [x0] + params)
  got = np.concatenate([_np(g).reshape(-1) for g in got])

  def f(x0v, pv):
    return float(np.sum(_oracle_paths(kind, pv, x0v, z) * w))
  want = []
  inputs = list(x00) + list(p0)
  for q in range(len(inputs)):
    h = 1e-6 * max(abs(inputs[q]), 1e-2)
    up, dn = list(inputs), list(inputs)
    up[q] += h
    dn[q] -= h
    want.append((f(up[:d], up[d:]) - f(dn[:d], dn[d:])) / (2 * h))
  want = np.array(want)
  scale = np.max(np.abs(want))
  np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6 * scale)


@pytest.mark.parametrize('name', ['gbm', 'heston_smooth'])
def test_tangent_vjp_equals_autograd_through_the_step_loop(name):
  # the generic loop is the reference's tape route: autograd over every
  # step's torch ops; same draws, same update order
  print('tangent vjp vs loop autograd', name, flush=True)
  kind, p0, x00, variant = CASES[name]
  d = len(x00)
  rs = np.random.RandomState(5)
  n = 4096
  z = rs.standard_normal((n, STEPS + 2, d))
  w = torch.as_tensor(rs.standard_normal((n, len(TIMES), d)), device='cuda')
  dev = 'cuda'

  def leaves():
    return ([torch.tensor(x00, dtype=torch.float64, device=dev,
                          requires_grad=True)] +
            [torch.tensor(v, dtype=torch.float64, device=dev,
                          requires_grad=True) for v in p0])
  a = leaves()
  fused = _fused(kind, a[1:], a[0], variant, z)
  ga = torch.autograd.grad((fused * w).sum(), a)
  b = leaves()
  if kind == 'gbm':
    mu, sig = b[1], b[2]
    drift = lambda t, x: mu * x
    vol = lambda t, x: sig * x.unsqueeze(-1)
  else:
    r, k, th, xi, rho = b[1:]

    def drift(t, x):
      vp = torch.clamp(x[..., 1], min=0)
      return torch.stack([r - vp / 2, k * (th - vp)], -1)

    def vol(t, x):
      vp = torch.sqrt(torch.clamp(x[..., 1], min=0))
      zz = torch.zeros_like(vp)
      return torch.stack([torch.stack([vp, xi * rho * vp], -1),
                          torch.stack([zz, xi * (1 - rho * rho)**0.5 * vp],
                                      -1)], -1)
  loop = tqf.models.euler_sampling.sample(
      d, drift, vol, TIMES, num_time_steps=STEPS, initial_state=b[0],
      normal_draws=z, watch_params=b[1:], dtype=torch.float64)
  np.testing.assert_allclose(_np(loop), _np(fused), rtol=1e-12)
  gb = torch.autograd.grad((loop * w).sum(), b)
  for u, v in zip(ga, gb):
    np.testing.assert_allclose(_np(u), _np(v), rtol=1e-9, atol=1e-12)


def test_fwd_gradient_through_fused_paths_is_the_tangent_jvp():
  # tqf.math.fwd_gradient's double reverse reaches the JVP kernel: the
  # tangent paths dX/dsigma equal central differences of the paths
  print('fwd_gradient through fused paths', flush=True)
  rs = np.random.RandomState(9)
  n = 3000
  z = rs.standard_normal((n, STEPS + 2, 2))
  p0 = CASES['heston'][1]
  x00 = CASES['heston'][2]

  def paths(xi):
    params = [torch.tensor(v, dtype=torch.float64, device='cuda')
              for v in p0]
    params[3] = xi
    return _fused('heston', params, x00, 0.0, z)
  xi = torch.tensor(p0[3], dtype=torch.float64, device='cuda')
  jvp = _np(tqf.math.fwd_gradient(paths, xi))
  h = 1e-7
  with torch.no_grad():
    fd = (_np(paths(xi + h)) - _np(paths(xi - h))) / (2 * h)
  assert jvp.shape == (n, len(TIMES), 2)
  scale = np.max(np.abs(fd))
  bad = np.abs(jvp - fd) > 1e-5 * scale
  # paths whose variance touches 0 inside the bump see the truncation kink
  assert bad.mean() < 1e-3, bad.mean()
  # and the reverse sum of the same tangents is the VJP with unit weights
  g = tqf.math.gradients(lambda v: paths(v).sum(), xi)
  np.testing.assert_allclose(float(g), jvp.sum(), rtol=1e-10)


def test_model_class_tensors_and_philox_stream():
  # GeometricBrownianMotion built from tensors that require grad: its
  # callables run fused and differentiate; Philox draws (regenerated by the
  # tangent kernel) give the pathwise vega of the fused forward's own
  # central differences with common random numbers
  print('model tensors + philox', flush=True)
  from tqf.math.random import multivariate_normal as mvn
  mu = torch.tensor(0.05, dtype=torch.float64, device='cuda')
  r, k = 0.05, 100.0

  def price(sig):
    gbm = tqf.models.GeometricBrownianMotion(mu, sig, dtype=torch.float64)
    p = tqf.models.euler_sampling.sample(
        1, gbm.drift_fn(), gbm.volatility_fn(), [1.0], num_time_steps=50,
        num_samples=2**18, initial_state=[100.0],
        random_type=mvn.RandomType.PSEUDO_ANTITHETIC, seed=7,
        dtype=torch.float64)
    return (torch.relu(p[:, -1, 0] - k) * np.exp(-r)).mean()
  sig = torch.tensor(0.2, dtype=torch.float64, device='cuda',
                     requires_grad=True)
  vega = float(torch.autograd.grad(price(sig), sig)[0])
  h = 1e-6
  with torch.no_grad():
    fd = (float(price(torch.tensor(0.2 + h, dtype=torch.float64,
                                   device='cuda'))) -
          float(price(torch.tensor(0.2 - h, dtype=torch.float64,
                                   device='cuda')))) / (2 * h)
  assert abs(vega - fd) <= 1e-4 * abs(fd), (vega, fd)
  from oracle import closed_form
  assert abs(vega - closed_form.bs_vega(100.0, k, 1.0, r, 0.2)) < 0.02 * fd


def test_per_path_initial_state_gradient():
  print('per-path x0 gradient', flush=True)
  rs = np.random.RandomState(2)
  n = 1000
  z = rs.standard_normal((n, STEPS + 2, 1))
  x0 = torch.as_tensor(100.0 + rs.uniform(-5, 5, (n, 1)), device='cuda')
  x0.requires_grad_(True)
  paths = _fused('gbm', [0.05, 0.25], x0, None, z)
  w = torch.as_tensor(rs.standard_normal((n, len(TIMES), 1)), device='cuda')
  g = _np(torch.autograd.grad((paths * w).sum(), x0)[0])
  # Euler on S is linear in S0 path by path: dX/dS0 = X / S0
  want = np.sum(_np(paths) * _np(w), axis=(1, 2)) / _np(x0)[:, 0]
  np.testing.assert_allclose(g[:, 0], want, rtol=1e-12)


def test_gradient_memory_is_flat_in_the_number_of_steps():
  # 2000 steps of 2^16 two-dimensional paths: the per-step loop would keep
  # gigabytes of intermediates for its tape; the tangent route keeps the
  # outputs and one workspace
  print('tangent memory', flush=True)
  from tqf.math.random import multivariate_normal as mvn
  torch.cuda.synchronize()
  torch.cuda.reset_peak_memory_stats()
  base = torch.cuda.memory_allocated()
  params = [torch.tensor(v, dtype=torch.float32, device='cuda',
                         requires_grad=True)
            for v in CASES['heston'][1]]
  spec = _spec('heston', params, 0.0)
  f = engine.tag(lambda t, x: None, spec)
  p = tqf.models.euler_sampling.sample(
      2, f, f, [1.0], num_time_steps=2000, num_samples=2**16,
      initial_state=CASES['heston'][2], random_type=mvn.RandomType.PSEUDO,
      seed=3, dtype=torch.float32)
  g = torch.autograd.grad(p[:, -1, 0].exp().mean(), params)
  torch.cuda.synchronize()
  peak = torch.cuda.max_memory_allocated() - base
  assert all(torch.isfinite(x) for x in g)
  assert peak < 64 * 2**20, peak
