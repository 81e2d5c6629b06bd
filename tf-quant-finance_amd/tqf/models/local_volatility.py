"""Local volatility model (mirror of
tff.experimental.local_volatility.LocalVolatilityModel).

  LocalVolatilityModel   tff/experimental/local_volatility/
                         local_volatility_model.py:307-540

The reference simulates X = ln S with the Euler scheme of GenericItoProcess
(local_volatility_model.py:461-476):
    drift(t, X) = r(t) - q(t) - sigma(t, e^X)^2 / 2
    vol(t, X)   = sigma(t, e^X)[..., None] * chol(corr)
and returns e^X.  Here the same composition runs in the fused Euler kernel
(`TQF_MODEL_LOCAL_VOL`: the surface staged in LDS, looked up bilinearly in
(t, ln S)) when the local volatility is a `LocalVolSurface` and the model is
one-dimensional with constant rates; any other local volatility callable,
rate callables or dim > 1 run as the per-step torch loop of
`euler_sampling.sample` on the same GPU with identical semantics.

`precompute_iv=True` with `times_grid` and `spot_grid` tabulates the given
callable once on that grid (the reference precomputes its spline surface in
the same situation, local_volatility_model.py:434-440, and then samples on
`times_grid`).
"""
import numpy as np
import torch

from tqf import _native
from tqf.engine import euler as engine
from tqf.models import euler_sampling
from tqf.models import utils

__all__ = ['LocalVolatilityModel', 'LocalVolSurface']


class LocalVolSurface:
  """sigma(t, S) tabulated on times t_i (increasing) and a uniform log-spot
  axis ln S_j = log_spot_min + j * log_spot_step; bilinear in (t, ln S),
  flat outside the table.  Callable as sigma(t, S) with S [N, 1] (the
  reference's local_volatility_fn convention) for the generic loop; the
  fused kernel reads the table itself."""

  def __init__(self, times, log_spot_min, log_spot_step, vols):
    self.times = np.asarray(times, np.float64).reshape(-1)
    self.log_spot_min = float(log_spot_min)
    self.log_spot_step = float(log_spot_step)
    self.vols = np.asarray(vols, np.float64)
    if self.vols.ndim != 2 or self.vols.shape[0] != self.times.shape[0]:
      raise ValueError('vols must have shape [len(times), n_spots]')
    if self.vols.shape[1] < 2 or self.log_spot_step <= 0:
      raise ValueError('need at least two spot nodes and a positive step')
    if np.any(np.diff(self.times) <= 0):
      raise ValueError('times must be increasing')

  @classmethod
  def from_function(cls, fn, times, spots, num_spots=None, dtype=np.float64):
    """Tabulates fn(t, S[N, 1]) -> [N, 1] at `times` on a uniform log grid
    spanning min(spots)..max(spots) with num_spots (default len(spots))
    nodes."""
    spots = np.asarray(spots, np.float64).reshape(-1)
    n = int(num_spots or spots.shape[0])
    x = np.linspace(np.log(spots.min()), np.log(spots.max()), n)
    s = np.exp(x)
    rows = []
    for t in np.asarray(times, np.float64).reshape(-1):
      v = fn(torch.tensor(t, dtype=torch.float64),
             torch.as_tensor(s.reshape(-1, 1), dtype=torch.float64))
      rows.append(np.asarray(utils.to_numpy(v), np.float64).reshape(-1))
    return cls(times, x[0], x[1] - x[0], np.stack(rows))

  def __call__(self, t, spot):
    """Device evaluation for the generic loop: same interpolation."""
    x = torch.log(spot)
    dev, dt = x.device, x.dtype
    nt, nx = self.vols.shape
    vols = torch.as_tensor(self.vols, dtype=dt, device=dev)
    times = torch.as_tensor(self.times, dtype=dt, device=dev)
    t = torch.as_tensor(t, dtype=dt, device=dev)
    fx = ((x - self.log_spot_min) / self.log_spot_step).clamp(0, nx - 1)
    j = torch.clamp(fx.floor().long(), max=nx - 2)
    wx = fx - j.to(dt)
    row = int(torch.searchsorted(times, t.reshape(1), right=True).item()) - 1
    row = max(row, 0)
    r1 = min(row + 1, nt - 1)
    wt = torch.zeros((), dtype=dt, device=dev)
    if row + 1 < nt and float(t) > self.times[row]:
      wt = (t - times[row]) / (times[row + 1] - times[row])
    v0, v1 = vols[row], vols[r1]
    s0 = v0[j] + wx * (v0[j + 1] - v0[j])
    s1 = v1[j] + wx * (v1[j + 1] - v1[j])
    return s0 + wt * (s1 - s0)


def _const_rate(v, dim):
  if v is None:
    return np.zeros(dim)
  if callable(v):
    return None
  return np.broadcast_to(np.asarray(utils.to_numpy(v), np.float64),
                         [dim]).copy()


class LocalVolatilityModel:
  """dS_i / S_i = (r - q) dt + sigma(t, S_i) dW_i, corr(dW) = corr_matrix."""

  def __init__(self, dim, risk_free_rate=None, dividend_yield=None,
               local_volatility_fn=None, corr_matrix=None, times_grid=None,
               spot_grid=None, precompute_iv=False, dtype=None, name=None):
    self._name = name or 'local_volatility_model'
    self._dim = int(dim)
    self._dtype = dtype if isinstance(dtype, torch.dtype) else (
        torch.float64 if dtype is None or np.dtype(dtype) == np.float64
        else torch.float32)
    self._times_grid = None
    self._precompute_iv = bool(precompute_iv)
    if local_volatility_fn is None:
      raise NotImplementedError(
          'local_volatility_fn is required (from_market_data / Dupire '
          'calibration is not part of this build)')
    if self._precompute_iv:
      if times_grid is None or spot_grid is None:
        raise ValueError(
            'When `precompute_iv` is True, both `times_grid` and `spot_grid` '
            'must be supplied')
      self._times_grid = np.asarray(utils.to_numpy(times_grid), np.float64)
      if not isinstance(local_volatility_fn, LocalVolSurface):
        local_volatility_fn = LocalVolSurface.from_function(
            local_volatility_fn, self._times_grid, utils.to_numpy(spot_grid))
    self._local_volatility_fn = local_volatility_fn
    self._rf = risk_free_rate
    self._q = dividend_yield
    self._corr = None if corr_matrix is None else np.asarray(
        utils.to_numpy(corr_matrix), np.float64)
    self._chol = (np.eye(self._dim) if self._corr is None
                  else np.linalg.cholesky(self._corr))

  def dim(self):
    return self._dim

  def dtype(self):
    return self._dtype

  def name(self):
    return self._name

  def local_volatility_fn(self):
    return self._local_volatility_fn

  def precompute_iv(self):
    return self._precompute_iv

  def _rate(self, v, t, like):
    if v is None:
      return torch.zeros(self._dim, dtype=like.dtype, device=like.device)
    if callable(v):
      return torch.as_tensor(v(t), dtype=like.dtype, device=like.device)
    return torch.as_tensor(utils.to_numpy(v), dtype=like.dtype,
                           device=like.device)

  def drift_fn(self):
    """drift(t, X) = r - q - sigma(t, e^X)^2 / 2 (ln S coordinates)."""
    lv = self._local_volatility_fn

    def _drift(t, x):
      s = lv(t, torch.exp(x))
      return self._rate(self._rf, t, x) - self._rate(self._q, t, x) - s * s / 2
    spec = self._spec()
    return engine.tag(_drift, spec) if spec else _drift

  def volatility_fn(self):
    """vol(t, X) = sigma(t, e^X)[..., None] * chol(corr)."""
    lv = self._local_volatility_fn
    chol = self._chol

    def _vol(t, x):
      s = lv(t, torch.exp(x))
      return s.unsqueeze(-1) * torch.as_tensor(chol, dtype=x.dtype,
                                               device=x.device)
    spec = self._spec()
    return engine.tag(_vol, spec) if spec else _vol

  def _spec(self):
    """('local_vol', table) when the fused kernel applies, else None."""
    if not hasattr(self, '_spec_cache'):
      spec = None
      r = _const_rate(self._rf, self._dim)
      q = _const_rate(self._q, self._dim)
      lv = self._local_volatility_fn
      if self._dim == 1 and r is not None and q is not None and isinstance(
          lv, LocalVolSurface):
        table = engine.LocalVolTable(lv.times, lv.log_spot_min,
                                     lv.log_spot_step, lv.vols,
                                     float(r[0] - q[0]))
        spec = ('local_vol', table)
      self._spec_cache = spec
    return self._spec_cache

  def sample_paths(self, times, num_samples=1, initial_state=None,
                   random_type=None, seed=None, swap_memory=True,
                   time_step=None, num_time_steps=None, skip=0,
                   precompute_normal_draws=True, times_grid=None,
                   normal_draws=None, watch_params=None,
                   validate_args=False, name=None):
    """Paths of S [num_samples, k, dim] (local_volatility_model.py:478-540).
    """
    del swap_memory, precompute_normal_draws, name
    if self._precompute_iv:
      if (time_step is not None or num_time_steps is not None or
          times_grid is not None):
        raise ValueError(
            '`time_step`, `num_time_steps`, or `times_grid` cannot be used'
            'with the interpolated LVM')
      times_grid = self._times_grid
    dev = _native.device()
    x0 = None
    if initial_state is not None:
      x0 = torch.log(torch.as_tensor(initial_state, device=dev).to(
          self._dtype))
    x = euler_sampling.sample(
        self._dim, self.drift_fn(), self.volatility_fn(), times,
        time_step=time_step, num_time_steps=num_time_steps,
        num_samples=num_samples, initial_state=x0, random_type=random_type,
        seed=seed, skip=skip, times_grid=times_grid,
        normal_draws=normal_draws, watch_params=watch_params,
        validate_args=validate_args, dtype=self._dtype)
    return torch.exp(x)
