"""Geometric Brownian motion (mirror of tff.models.GeometricBrownianMotion).

  GeometricBrownianMotion             univariate_geometric_brownian_motion.py:
                                      26-340 (constant mean/volatility)
  MultivariateGeometricBrownianMotion multivariate_geometric_brownian_motion.py:
                                      25-282

`drift_fn()` / `volatility_fn()` return the reference's coefficient
callables (usable with any Euler sampler); they carry a spec so
`tqf.models.euler_sampling.sample` runs them in the fused Euler kernel.
`sample_paths` is the reference's exact log-normal scheme, run by the
`tqf_gbm_exact` kernel with the normals generated in place.
"""
import ctypes

import numpy as np
import torch

from tqf import _native
from tqf.engine import euler as engine
from tqf.math.random import sobol as sobol_lib
from tqf.models import utils

__all__ = ['GeometricBrownianMotion', 'MultivariateGeometricBrownianMotion',
           'run_exact']


def _dtype_of(dtype, *vals):
  if dtype is not None:
    if isinstance(dtype, torch.dtype):
      return dtype
    return torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
  for v in vals:
    if isinstance(v, torch.Tensor) and v.is_floating_point():
      return v.dtype
    if isinstance(v, np.ndarray) and v.dtype == np.float64:
      return torch.float64
  return torch.float32


def run_exact(times, drift, vol, chol, s0, num_samples, rng, dtype,
              keep_mask=None, mode='paths', dev=None):
  """Runs tqf_gbm_exact over the grid [0] + times (every time recorded
  unless keep_mask says otherwise).  mode 'paths': [N, k, d] contiguous;
  'store': [k, d, N]; 'view': the [N, k, d] view of the [k, d, N] store."""
  dev = dev or _native.device()
  npd = utils.np_dtype(dtype)
  times = np.asarray(times, npd).reshape(-1)
  all_times = np.concatenate([np.zeros(1, npd), times]).astype(npd)
  if keep_mask is None:
    keep_mask = np.concatenate([[False], np.ones(times.shape[0], bool)])
  k = int(np.sum(keep_mask[1:])) + int(keep_mask[0])
  dt = (all_times[1:] - all_times[:-1]).astype(npd)
  sqrt_dt = np.sqrt(dt).astype(npd)
  slots, init_slot, executed, _ = utils.record_plan(keep_mask, k)
  d = int(np.asarray(drift).shape[0])
  keep = []
  # the small arrays go up in one staged copy (addresses set by flush)
  stage = engine.Staging(dev)

  g = _native.Gbm()
  g.dim = d
  stage.add(utils.to_numpy(drift, np.float64), dtype, [(g, 'drift')])
  stage.add(utils.to_numpy(vol, np.float64), dtype, [(g, 'vol')])
  if chol is None:
    g.chol = 0
  else:
    stage.add(utils.to_numpy(chol, np.float64), dtype, [(g, 'chol')])
  if not torch.is_tensor(s0) and np.ndim(s0) <= 1:
    # one state for every path
    s0_h = np.asarray(utils.to_numpy(s0), utils.np_dtype(dtype)).reshape(-1)
    stage.add(np.broadcast_to(s0_h, (d,)), dtype, [(g, 's0')])
    g.s0_stride_path = 0
  else:
    s0_t = utils.device_tensor(s0, dtype, dev)
    if s0_t.dim() == 0:
      s0_t = s0_t.reshape(1)
    if s0_t.dim() == 1:
      s0_t = s0_t.expand(d).contiguous()
      g.s0_stride_path = 0
    else:
      s0_t = s0_t.expand(num_samples, d).contiguous()
      g.s0_stride_path = d
    keep.append(s0_t)
    g.s0 = s0_t.data_ptr()

  grid = _native.Grid()
  stage.add(dt[:max(executed, 1)], dtype, [(grid, 'dt')])
  stage.add(sqrt_dt[:max(executed, 1)], dtype, [(grid, 'sqrt_dt')])
  stage.add(np.concatenate([slots, [-1]]).astype(np.int32), torch.int32,
            [(grid, 'slot')])
  grid.n_steps, grid.n_slots, grid.init_slot = executed, k, init_slot
  grid.n_rng_steps = dt.shape[0]

  r = _native.Rng()
  r.kind = engine._RNG_CODES[rng.kind]
  r.seed = int(rng.seed) & 0xFFFFFFFFFFFFFFFF
  r.stream_id = int(rng.stream_id)
  if rng.kind == 'draws':
    draws = rng.draws.to(device=dev, dtype=dtype).contiguous()
    if draws.dim() != 3 or draws.shape[2] != d:
      raise ValueError('`dim` should be equal to `normal_draws.shape[2]` '
                       'but are {0} and {1} respectively'.format(
                           d, draws.shape[-1]))
    if draws.shape[1] < executed:
      raise ValueError('normal_draws must have shape [num_samples, '
                       'num_time_points, dim]')
    keep.append(draws)
    r.draws = draws.data_ptr()
    r.stride_path, r.stride_step = draws.stride(0), draws.stride(1)
  elif rng.kind in ('sobol', 'sobol_shifted'):
    ndims = max(dt.shape[0], 1) * d
    nd = (sobol_lib.num_digits_for(rng.skip, num_samples)
          if rng.kind == 'sobol' else 32)
    r.sobol_v = sobol_lib.device_columns(ndims, nd, dev=dev).data_ptr()
    r.sobol_lo = sobol_lib.device_lo_table(ndims, nd, group=d,
                                           dev=dev).data_ptr()
    r.sobol_first_index = rng.skip + 1 if rng.kind == 'sobol' else rng.first_index
    r.sobol_nd, r.sobol_ndims = nd, ndims
    if rng.shift is not None:
      stage.add(np.asarray(rng.shift, np.uint64).astype(np.uint32).view(
          np.int32), torch.int32, [(r, 'sobol_shift')])
  # every slot is written unless a requested time lies past the executed
  # steps; only then does the output need zeros first
  written = np.zeros(k, bool)
  done = np.asarray(slots[:executed])
  written[done[done >= 0]] = True
  if init_slot >= 0:
    written[init_slot] = True
  alloc = torch.empty if written.all() else torch.zeros
  if mode == 'paths':
    out = alloc((num_samples, k, d), dtype=dtype, device=dev)
    code = _native.OUT_PATH_MAJOR
  else:
    out = alloc((k, d, num_samples), dtype=dtype, device=dev)
    code = _native.OUT_STEP_MAJOR
  stage.flush(keep)
  _native.call('tqf_gbm_exact', ctypes.byref(g), ctypes.byref(grid),
               ctypes.byref(r), ctypes.c_int32(code),
               ctypes.c_void_p(out.data_ptr()), ctypes.c_int64(num_samples),
               ctypes.c_int32(_native.dtype_code(dtype)),
               _native.stream_ptr(dev))
  if mode == 'view':
    out = out.permute(2, 0, 1)
  out._tqf_keep = keep
  return out


class GeometricBrownianMotion:
  """dX = mu X dt + sigma X dW with constant mu, sigma."""

  def __init__(self, mean, volatility, dtype=None, name=None):
    self._name = name or 'geometric_brownian_motion'
    self._dtype = _dtype_of(dtype, mean, volatility)
    if np.ndim(utils.to_numpy(mean)) or np.ndim(utils.to_numpy(volatility)):
      raise NotImplementedError(
          'batched or piecewise-constant GBM parameters are not supported')
    # tensors that require grad are kept: euler_sampling.sample then
    # differentiates with respect to them
    self._mean = utils.scalar_param(mean)
    self._volatility = utils.scalar_param(volatility)

  def dim(self):
    return 1

  def dtype(self):
    return self._dtype

  def name(self):
    return self._name

  def drift_fn(self):
    mu = self._mean

    def _drift(t, x):
      del t
      return mu * x
    return engine.tag(_drift, ('gbm', mu, self._volatility))

  def volatility_fn(self):
    sigma = self._volatility

    def _vol(t, x):
      del t
      return sigma * x.unsqueeze(-1)
    return engine.tag(_vol, ('gbm', self._mean, sigma))

  def sample_paths(self, times, initial_state=None, num_samples=1,
                   random_type=None, seed=None, skip=0, normal_draws=None,
                   name=None):
    """Exact log-normal paths [num_samples, k, 1] at `times` (a view of
    the coalesced [k, 1, num_samples] store)."""
    del name
    dev = _native.device()
    times_np = utils.to_numpy(times, utils.np_dtype(self._dtype)).reshape(-1)
    if initial_state is None:
      initial_state = 1.0
    s0 = utils.device_tensor(initial_state, self._dtype, dev)
    if s0.dim() > 1 or (s0.dim() == 1 and s0.shape[0] not in (1,)):
      s0 = s0.reshape(-1, 1)
    draws = None
    if normal_draws is not None:
      draws = utils.device_tensor(normal_draws, self._dtype, dev)
      if draws.dim() != 3 or draws.shape[2] != 1:
        raise ValueError('`dim` should be equal to `1` but is {0}'.format(
            draws.shape[-1]))
      num_samples = draws.shape[0]
    rng = engine.RngSpec.from_random_type(random_type, seed, skip, draws)
    mu, sig = float(self._mean), float(self._volatility)

    def run():
      return run_exact(times_np, [mu - sig * sig / 2], [sig], None,
                       s0.detach(), int(num_samples), rng, self._dtype,
                       mode='view', dev=dev)
    params = [self._mean, self._volatility]
    if torch.is_grad_enabled() and (s0.requires_grad or any(
        torch.is_tensor(p) and p.requires_grad for p in params)):
      t = torch.as_tensor(times_np, dtype=self._dtype, device=dev)
      mu_t, sig_t = (p if torch.is_tensor(p) else
                     torch.tensor(p, dtype=self._dtype, device=dev)
                     for p in params)
      return _ExactPaths.apply(run, t, s0, mu_t, sig_t)
    return run()


class _ExactPaths(torch.autograd.Function):
  """Exact GBM paths S = S0 exp((mu - sigma^2/2) t + sigma W_t) from
  tqf_gbm_exact, differentiable in S0, mu and sigma.  The backward needs
  only the paths: dS/dS0 = S / S0, dS/dmu = S t, dS/dsigma = S (W_t -
  sigma t) with W_t = (ln(S / S0) - (mu - sigma^2/2) t) / sigma; at
  sigma = 0 the sigma-derivative is 0, the reference's convention for
  sqrt(sigma^2 dt) (univariate_geometric_brownian_motion.py:712-718)."""

  @staticmethod
  def forward(ctx, run, t, s0, mu, sig):
    out = run()
    ctx.save_for_backward(out, t, s0, mu, sig)
    return out

  @staticmethod
  def backward(ctx, g):
    out, t, s0, mu, sig = ctx.saved_tensors
    n = out.shape[0]
    s0b = s0.reshape(-1, 1, 1) if s0.numel() > 1 else s0.reshape(1, 1, 1)
    gs = g * out
    tt = t.reshape(1, -1, 1)
    g_s0 = (gs / s0b).sum(dim=(1, 2))
    g_s0 = (g_s0.reshape(s0.shape) if s0.numel() == n and s0.numel() > 1
            else g_s0.sum().reshape(s0.shape))
    g_mu = (gs * tt).sum().reshape(mu.shape)
    if float(sig) != 0.0:
      w = (torch.log(out / s0b) - (mu - 0.5 * sig * sig) * tt) / sig
      g_sig = (gs * (w - sig * tt)).sum().reshape(sig.shape)
    else:
      g_sig = torch.zeros_like(sig)
    return None, None, g_s0, g_mu, g_sig


class MultivariateGeometricBrownianMotion:
  """Correlated GBM: dX_i = mu_i X_i dt + sigma_i X_i dW_i, corr(W)."""

  def __init__(self, dim, means=0.0, volatilities=1.0, corr_matrix=None,
               dtype=None, name=None):
    self._name = name or 'multivariate_geometric_brownian_motion'
    self._dim = int(dim)
    self._dtype = _dtype_of(dtype, means, volatilities)
    self._means = np.broadcast_to(np.asarray(utils.to_numpy(means), np.float64),
                                  [self._dim]).copy()
    self._vols = np.broadcast_to(
        np.asarray(utils.to_numpy(volatilities), np.float64), [self._dim]).copy()
    if corr_matrix is None:
      self._corr = None
    else:
      c = np.asarray(utils.to_numpy(corr_matrix), np.float64)
      if list(c.shape) != [self._dim, self._dim]:
        raise ValueError('`corr_matrix` must be of shape [{0}, {0}] but is '
                         'of shape {1}'.format(self._dim, list(c.shape)))
      self._corr = c

  def dim(self):
    return self._dim

  def dtype(self):
    return self._dtype

  def name(self):
    return self._name

  def cholesky(self):
    return None if self._corr is None else np.linalg.cholesky(self._corr)

  def drift_fn(self):
    means = self._means

    def _drift(t, x):
      del t
      return torch.as_tensor(means, dtype=x.dtype, device=x.device) * x
    return _drift

  def volatility_fn(self):
    vols, chol = self._vols, self.cholesky()

    def _vol(t, x):
      del t
      v = torch.as_tensor(vols, dtype=x.dtype, device=x.device) * x
      if chol is None:
        return torch.diag_embed(v)
      return v.unsqueeze(-1) * torch.as_tensor(chol, dtype=x.dtype,
                                               device=x.device)
    return _vol

  def sample_paths(self, times, initial_state=None, num_samples=1,
                   random_type=None, seed=None, skip=0, normal_draws=None,
                   name=None):
    """Exact correlated log-normal paths [num_samples, k, dim]."""
    del name
    dev = _native.device()
    times_np = utils.to_numpy(times, utils.np_dtype(self._dtype)).reshape(-1)
    if initial_state is None:
      initial_state = np.ones(self._dim)
    s0 = utils.device_tensor(initial_state, self._dtype, dev)
    draws = None
    if normal_draws is not None:
      draws = utils.device_tensor(normal_draws, self._dtype, dev)
      if draws.dim() != 3 or draws.shape[2] != self._dim:
        raise ValueError(
            '`dim` should be equal to `normal_draws.shape[2]` but are '
            '{0} and {1} respectively'.format(self._dim, draws.shape[-1]))
      num_samples = draws.shape[0]
    rng = engine.RngSpec.from_random_type(random_type, seed, skip, draws)
    drift = self._means - self._vols**2 / 2
    return run_exact(times_np, drift, self._vols, self.cholesky(), s0,
                     int(num_samples), rng, self._dtype, mode='view', dev=dev)
