"""Heston stochastic volatility model (mirror of tff.models.HestonModel).

  dX = -V/2 dt + sqrt(V) dW_X,  dV = kappa (theta - V) dt + xi sqrt(V) dW_V,
  corr(dW_X, dW_V) = rho                  (tff/models/heston/heston_model.py)

  drift_fn / volatility_fn   heston_model.py:143-173 (raw v in the drift,
                             sqrt|v| in the volatility); with constant
                             parameters they run in the fused Euler kernel
                             (variant 1 of TQF_MODEL_HESTON_FT).
  sample_paths               heston_model.py:177-460, Andersen QE with
                             psi_c = 1.5 and gamma_1 = gamma_2 = 1/2, fused
                             in the TQF_MODEL_HESTON_QE kernel: one thread per
                             path, per-step constants from the host.
  expected_total_variance    heston_model.py:462-508
  euler_ft_fns(r)            engine extension: the full-truncation log-Euler
                             coefficients [r - v+/2, kappa (theta - v+)] and
                             [[sqrt v+, 0], [xi rho sqrt v+, xi sqrt(1-rho^2)
                             sqrt v+]] for euler_sampling.sample.
"""
import numpy as np
import torch

from tqf import _native
from tqf.engine import euler as engine
from tqf.math.piecewise import PiecewiseConstantFunc
from tqf.math.random.multivariate_normal import RandomType
from tqf.models import utils

__all__ = ['HestonModel', 'qe_step_coefficients']


def _param_at(p, t):
  if isinstance(p, PiecewiseConstantFunc):
    return p(t)
  return np.full(np.shape(t), float(p))


def qe_step_coefficients(all_times, kappa, theta, xi, rho, tolerance, dtype):
  """[steps, 12] per-step QE constants (layout of tqf_model.step_coef).

  Parameters are evaluated at times + min(dt)/2 (heston_model.py:338-342).
  """
  t = np.asarray(all_times, np.float64)
  dt = t[1:] - t[:-1]
  mid = t[:-1] + (dt.min() / 2 if dt.size else 0.0)
  k, th, x, r = (_param_at(p, mid) for p in (kappa, theta, xi, rho))
  e = np.exp(-k * dt)
  c = np.zeros((dt.shape[0], 12))
  c[:, 0] = e
  c[:, 1] = x * x * e / k * (1 - e)
  c[:, 2] = th * x * x / 2 / k * (1 - e)**2
  c[:, 3] = -r * k * th / x * dt
  c[:, 4] = 0.5 * dt * (k * r / x - 0.5) - r / x
  c[:, 5] = 0.5 * dt * (k * r / x - 0.5) + r / x
  c[:, 6] = 0.5 * dt * (1 - r * r)
  c[:, 7] = 0.5 * dt * (1 - r * r)
  c[:, 8] = th
  c[:, 9] = (dt.astype(utils.np_dtype(dtype)) > tolerance).astype(np.float64)
  return c


class HestonModel:
  """Heston model with scalar or piecewise-constant parameters."""

  def __init__(self, mean_reversion, theta, volvol, rho, dtype=None,
               name=None):
    self._name = name or 'heston_model'
    if dtype is None:
      dtype = torch.float32
    elif not isinstance(dtype, torch.dtype):
      dtype = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    self._dtype = dtype

    def conv(p):
      if isinstance(p, PiecewiseConstantFunc):
        return p
      return utils.scalar_param(p)
    self._mean_reversion = conv(mean_reversion)
    self._theta = conv(theta)
    self._volvol = conv(volvol)
    self._rho = conv(rho)
    self._params = (self._mean_reversion, self._theta, self._volvol,
                    self._rho)
    self._constant = not any(isinstance(p, PiecewiseConstantFunc)
                             for p in self._params)

  def dim(self):
    return 2

  def dtype(self):
    return self._dtype

  def name(self):
    return self._name

  def _spec(self, r, variant):
    if not self._constant:
      return None
    return ('heston', r, self._mean_reversion, self._theta, self._volvol,
            self._rho, float(variant))

  def drift_fn(self):
    params = self._params

    def _drift(t, x):
      k, th = (torch.as_tensor(_param_at(p, utils.to_numpy(t)), dtype=x.dtype,
                               device=x.device) for p in params[:2])
      var = x[..., 1]
      return torch.stack([-var / 2, k * (th - var)], -1)
    spec = self._spec(0.0, 1)
    return engine.tag(_drift, spec) if spec else _drift

  def volatility_fn(self):
    params = self._params

    def _vol(t, x):
      xi, rho = (torch.as_tensor(_param_at(p, utils.to_numpy(t)),
                                 dtype=x.dtype, device=x.device)
                 for p in params[2:])
      vol = torch.sqrt(torch.abs(x[..., 1]))
      zeros = torch.zeros_like(vol)
      c1 = torch.stack([vol, xi * rho * vol], -1)
      c2 = torch.stack([zeros, xi * torch.sqrt(1 - rho**2) * vol], -1)
      return torch.stack([c1, c2], -1)
    spec = self._spec(0.0, 1)
    return engine.tag(_vol, spec) if spec else _vol

  def euler_ft_fns(self, risk_free_rate=0.0):
    """(drift_fn, volatility_fn) of the full-truncation log-Euler scheme."""
    if not self._constant:
      raise NotImplementedError('full truncation needs constant parameters')
    r = utils.scalar_param(risk_free_rate)
    k, th, xi, rho = self._params
    s1r = (1 - rho * rho)**0.5

    def _drift(t, x):
      vp = torch.clamp(x[..., 1], min=0)
      return torch.stack([r - vp / 2, k * (th - vp)], -1)

    def _vol(t, x):
      vp = torch.sqrt(torch.clamp(x[..., 1], min=0))
      z = torch.zeros_like(vp)
      c1 = torch.stack([vp, xi * rho * vp], -1)
      c2 = torch.stack([z, xi * s1r * vp], -1)
      return torch.stack([c1, c2], -1)
    spec = self._spec(r, 0)
    return engine.tag(_drift, spec), engine.tag(_vol, spec)

  def sample_paths(self, times, initial_state, num_samples=1,
                   random_type=None, seed=None, time_step=None, skip=0,
                   tolerance=1e-6, num_time_steps=None,
                   precompute_normal_draws=True, times_grid=None,
                   normal_draws=None, name=None):
    """QE paths [num_samples, k, 2] of (log spot, variance) at `times`."""
    del precompute_normal_draws, name
    if random_type is None:
      random_type = RandomType.PSEUDO
    dtype = self._dtype
    npd = utils.np_dtype(dtype)
    dev = _native.device()
    times_np = utils.to_numpy(times, npd).reshape(-1)
    draws = None
    if normal_draws is not None:
      draws = utils.device_tensor(normal_draws, dtype, dev)
      num_samples = draws.shape[0]
    if times_grid is None:
      if time_step is None:
        if num_time_steps is None:
          raise ValueError(
              'When `times_grid` is not supplied, either `num_time_steps` '
              'or `time_step` should be defined.')
        time_step = times_np[-1] / npd(num_time_steps)
      elif num_time_steps is not None:
        raise ValueError(
            'Both `time_step` and `num_time_steps` can not be `None` '
            'simultaneously when calling sample_paths of HestonModel.')
      extra = [p.jump_locations() for p in self._params
               if isinstance(p, PiecewiseConstantFunc)]
      all_times, keep_mask = utils.heston_grid(times_np, time_step,
                                               num_time_steps, extra, npd)
    else:
      all_times, keep_mask, _ = utils.prepare_grid(
          times=times_np, time_step=time_step, times_grid=times_grid,
          dtype=npd)
    coef = qe_step_coefficients(all_times, *self._params,
                                tolerance=tolerance, dtype=dtype)
    x0 = utils.to_numpy(initial_state, np.float64).reshape(-1)
    rng = engine.RngSpec.from_random_type(random_type, seed, skip, draws)
    spec = ('heston_qe', coef)
    return engine.run(spec, all_times, keep_mask, times_np.shape[0], x0,
                      int(num_samples), rng, dtype, mode='view',
                      initial_written=0, dev=dev)

  def expected_total_variance(self, future_times, initial_var, name=None):
    """E[int_0^T V dt] (heston_model.py:462-508); constant parameters only."""
    del name
    if not self._constant:
      raise ValueError('Expected total variance is not supported for '
                       'piecewise-constant parameters')
    k, th = self._mean_reversion, self._theta
    t = np.asarray(utils.to_numpy(future_times), np.float64)
    v0 = np.asarray(utils.to_numpy(initial_var), np.float64)
    strength = (1 - np.exp(-k * t)) / k
    return torch.as_tensor((v0 - th) * strength + th * t)
