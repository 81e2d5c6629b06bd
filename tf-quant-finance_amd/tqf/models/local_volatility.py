"""Local volatility model (mirror of
tff.experimental.local_volatility.LocalVolatilityModel).

  LocalVolatilityModel      tff/experimental/local_volatility/
                            local_volatility_model.py:307-540
  from_market_data          :542-696
  from_volatility_surface   :698-800
  Dupire local volatility   :38-99 (option prices), :102-141 (implied
                            volatilities), :144-304 (precomputed splines)

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

Dupire surfaces built by `from_market_data` / `from_volatility_surface`
evaluate the reference's formulas with torch forward-mode derivatives
(torch.func.jvp in place of tff.math.fwd_gradient), per step in the
generic loop.  With `precompute_iv=True` the reference freezes the
implied-volatility splines on `times_grid` rows and steps on `times_grid`;
there the local volatility at each grid time is a fixed function of S, so
for a one-dimensional model with constant rates it is tabulated once, at
the exact grid times, on a fine uniform ln S axis (`num_spot_nodes`,
default 2049) and the fused kernel reads that table: exact in t, linear in
ln S between nodes (the only difference from the per-step formula; flat
outside spot_grid, where the reference's value is constant in S too).
"""
import functools

import numpy as np
import torch

from tqf import _native
from tqf.engine import euler as engine
from tqf.experimental import volatility_surface as vs_lib
from tqf.math.interpolation import cubic
from tqf.math.interpolation import interpolation_2d
from tqf.math.interpolation._util import as_tensor, divide_no_nan
from tqf.models import euler_sampling
from tqf.models import utils

__all__ = ['LocalVolatilityModel', 'LocalVolSurface']


class LocalVolSurface:
  """sigma(t, S) tabulated on times t_i (increasing) and a uniform log-spot
  axis ln S_j = log_spot_min + j * log_spot_step; bilinear in (t, ln S),
  flat outside the table.  Callable as sigma(t, S) with S [N, 1] (the
  reference's local_volatility_fn convention) for the generic loop; the
  fused kernel reads the table itself."""

  def __init__(self, times, log_spot_min, log_spot_step, vols,
               outside=None):
    self.times = np.asarray(times, np.float64).reshape(-1)
    self.log_spot_min = float(log_spot_min)
    self.log_spot_step = float(log_spot_step)
    self.vols = np.asarray(vols, np.float64)
    # [nt, 2]: values below the first / above the last node (else flat)
    self.outside = None if outside is None else np.asarray(outside,
                                                           np.float64)
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
    fraw = (x - self.log_spot_min) / self.log_spot_step
    fx = fraw.clamp(0, nx - 1)
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
    if self.outside is not None:
      o = torch.as_tensor(self.outside, dtype=dt, device=dev)
      lo, hi = fraw < 0, fraw > nx - 1
      s0 = torch.where(lo, o[row, 0], torch.where(hi, o[row, 1], s0))
      s1 = torch.where(lo, o[r1, 0], torch.where(hi, o[r1, 1], s1))
    return s0 + wt * (s1 - s0)


def _fwd_gradient(fn, x):
  """d fn / d x by forward-mode AD with a unit tangent (elementwise
  derivative for elementwise fn; d/dt of every output for a scalar t) --
  tff.math.fwd_gradient."""
  x = torch.as_tensor(x)
  return torch.func.jvp(fn, (x,), (torch.ones_like(x),))[1]


def _rate_from_discount_factor(discount_factor_fn):
  """r(t) = d/dt (-ln P(t)) (local_volatility_model.py:814-822)."""

  def risk_free_rate_fn(t):
    # evaluated on the host: discount curves are small host-side callables
    t = torch.as_tensor(t)
    dev = t.device
    tc = t.to('cpu', torch.float64)
    r = _fwd_gradient(
        lambda x: -torch.log(torch.as_tensor(discount_factor_fn(x))), tc)
    return r.to(dev)
  return risk_free_rate_fn


def _ncdf(x):
  return (torch.erf(x / np.sqrt(2.0)) + 1) / 2


def _bs_call(vols, strikes, expiries, spots, dividend_rates,
             discount_factors):
  """Black-Scholes call prices in the reference's forward form
  (tff/black_scholes/vanilla_prices.py:133-183, lognormal, calls)."""
  discount_rates = -torch.log(discount_factors) / expiries
  forwards = spots * torch.exp((discount_rates - dividend_rates) * expiries)
  sqrt_var = vols * torch.sqrt(expiries)
  d1 = divide_no_nan(torch.log(forwards / strikes), sqrt_var) + sqrt_var / 2
  d2 = d1 - sqrt_var
  calls = torch.where(sqrt_var > 0,
                      forwards * _ncdf(d1) - strikes * _ncdf(d2),
                      torch.clamp(forwards - strikes, min=0.0))
  return discount_factors * calls


def _dupire_local_volatility_prices(time, spot_price, initial_spot_price,
                                    implied_volatility_surface,
                                    discount_factor_fn, dividend_yield):
  """Dupire from call prices C(T, K) of the implied surface
  (local_volatility_model.py:38-99):
  sigma^2 = 2 (C_T + (r - q) K C_K + q C) / (K^2 C_KK) at K = S, T = t."""
  time = as_tensor(time, like=spot_price)
  dtype = spot_price.dtype
  r = torch.as_tensor(_rate_from_discount_factor(discount_factor_fn)(time),
                      dtype=dtype, device=spot_price.device)
  q = as_tensor(dividend_yield, dtype, like=spot_price)
  s0 = as_tensor(initial_spot_price, dtype, like=spot_price)

  def _option_price(expiry_time, strike):
    dfs = torch.as_tensor(discount_factor_fn(expiry_time), dtype=dtype,
                          device=spot_price.device)
    vols = implied_volatility_surface(strike=strike,
                                      expiry_times=expiry_time)
    return _bs_call(vols, strike, expiry_time, s0, q, dfs)

  dcdk_fn = lambda x: _option_price(time, x)
  dcdt_fn = lambda x: _option_price(x, spot_price)
  d2cdk2_fn = lambda x: _fwd_gradient(dcdk_fn, x)
  numerator = (_fwd_gradient(dcdt_fn, time) +
               (r - q) * spot_price * _fwd_gradient(dcdk_fn, spot_price) +
               q * _option_price(time, spot_price))
  denominator = _fwd_gradient(d2cdk2_fn, spot_price) * spot_price**2
  var = torch.relu(2 * divide_no_nan(numerator, denominator))
  return torch.sqrt(var)


def _dupire_local_volatility_iv(time, spot_price, initial_spot_price,
                                implied_volatility_surface,
                                discount_factor_fn, dividend_yield):
  """Dupire in implied-volatility form (local_volatility_model.py:102-141),
  theta = sigma_imp(t, S) and its t, S, SS derivatives."""
  time = as_tensor(time, like=spot_price)
  dtype = spot_price.dtype
  r = torch.as_tensor(_rate_from_discount_factor(discount_factor_fn)(time),
                      dtype=dtype, device=spot_price.device)
  q = as_tensor(dividend_yield, dtype, like=spot_price)
  s0 = as_tensor(initial_spot_price, dtype, like=spot_price)

  def _iv(expiry_time, strike):
    return implied_volatility_surface(strike=strike, expiry_times=expiry_time)

  theta = _iv(time, spot_price)
  d1 = divide_no_nan(
      (torch.log(s0 / spot_price) + (r - q + 0.5 * theta**2) * time) / theta,
      torch.sqrt(time))
  spot_fn = lambda x: _iv(time, x)
  time_fn = lambda t: _iv(t, spot_price)
  dtheta_dt = _fwd_gradient(time_fn, time)
  dtheta_ds = _fwd_gradient(spot_fn, spot_price)
  d2theta_ds2 = _fwd_gradient(lambda x: _fwd_gradient(spot_fn, x),
                              spot_price)
  numerator = (theta**2 + 2 * time * theta * dtheta_dt +
               2 * (r - q) * spot_price * time * theta * dtheta_ds)
  denominator = ((1 + spot_price * d1 * time * dtheta_ds)**2 +
                 spot_price**2 * time * theta *
                 (d2theta_ds2 - d1 * time * dtheta_ds**2))
  return torch.sqrt(torch.relu(divide_no_nan(numerator, denominator)))


class _PrecomputedDupire:
  """The reference's precomputed Dupire function
  (local_volatility_model.py:144-304): implied vols and their time
  derivatives on times_grid x spot_grid, natural splines in S per row; at
  time t the row of times_grid[k] with t in (times_grid[k],
  times_grid[k+1]] (left-continuous on times_grid[1:]) is used.  Callable
  as f(t, S[N, dim]); `table()` tabulates it for the fused kernel."""

  def __init__(self, initial_spot_price, implied_volatility_surface,
               discount_factor_fn, dividend_yield, times_grid, spot_grid,
               dim, dtype):
    self._dtype = dtype
    self._dim = dim
    self._df = discount_factor_fn
    self._rate = _rate_from_discount_factor(discount_factor_fn)
    self._q = as_tensor(dividend_yield, dtype)
    self._s0 = as_tensor(initial_spot_price, dtype)
    self._tg = as_tensor(times_grid, dtype).reshape(-1)
    self._sg = as_tensor(spot_grid, dtype).reshape(-1)
    spots = self._sg.reshape(-1, 1)
    ivs, grads = [], []
    for t in self._tg:
      f = lambda tt: implied_volatility_surface(strike=spots,
                                                expiry_times=tt)
      ivs.append(f(t))
      grads.append(_fwd_gradient(f, t))
    nt, ns = self._tg.shape[0], self._sg.shape[0]
    # [nt, dim, ns]
    self._iv = torch.stack(ivs).reshape(nt, ns, -1).transpose(1, 2)
    self._grad = torch.stack(grads).reshape(nt, ns, -1).transpose(1, 2)
    nd = self._iv.shape[1]
    self._spot2d = self._sg.expand(nt, nd, ns)
    self._spline = cubic.build_spline(self._spot2d, self._iv)
    self._grad_spline = cubic.build_spline(self._spot2d, self._grad)

  def times_grid(self):
    return self._tg

  def _row(self, time):
    jumps = self._tg[1:]
    k = int(torch.searchsorted(jumps, time.detach().reshape(1).to(
        jumps.device), right=False).item())
    sp = cubic.SplineParameters(self._spot2d[k], self._iv[k],
                                self._spline.spline_coeffs[k])
    gp = cubic.SplineParameters(self._spot2d[k], self._grad[k],
                                self._grad_spline.spline_coeffs[k])
    return sp, gp

  def __call__(self, time, spot_price):
    dtype = spot_price.dtype
    dev = spot_price.device
    time = as_tensor(time, dtype, like=spot_price)
    r = torch.as_tensor(self._rate(time), dtype=dtype, device=dev)
    q = self._q.to(dev)
    s0 = self._s0.to(dev)
    sp, gp = self._row(time)
    sp = cubic.SplineParameters(*(a.to(dev) for a in (
        sp.x_data, sp.y_data, sp.spline_coeffs)))
    gp = cubic.SplineParameters(*(a.to(dev) for a in (
        gp.x_data, gp.y_data, gp.spline_coeffs)))
    theta_fn = lambda s: cubic.interpolate(s.transpose(0, 1), sp).transpose(
        0, 1)
    theta = theta_fn(spot_price)
    grad_times = cubic.interpolate(spot_price.transpose(0, 1),
                                   gp).transpose(0, 1)
    d1 = divide_no_nan(
        (torch.log(divide_no_nan(s0, spot_price)) +
         (r - q + 0.5 * theta**2) * time) / theta, torch.sqrt(time))
    grad_spots_fn = lambda s: _fwd_gradient(theta_fn, s)
    grad_spots = grad_spots_fn(spot_price)
    grad_grad_spots = _fwd_gradient(grad_spots_fn, spot_price)
    local_var = ((theta**2 + 2 * time * theta * grad_times +
                  2 * (r - q) * spot_price * time * theta * grad_spots) /
                 ((1 + spot_price * d1 * time * grad_spots)**2 +
                  spot_price**2 * time * theta *
                  (grad_grad_spots - d1 * time * grad_spots**2)))
    return torch.sqrt(torch.relu(local_var))

  def constant_rate(self):
    """r - q as a float when r(t) is the same at every grid time, else
    None."""
    r = np.array([float(torch.as_tensor(self._rate(t)).reshape(-1)[0])
                  for t in self._tg.cpu()])
    if np.ptp(r) > 1e-12 * max(1.0, np.abs(r).max()):
      return None
    q = self._q.reshape(-1).cpu().numpy()
    if q.shape[0] > 1 and np.ptp(q) > 0:
      return None
    return float(r[0] - q[0])

  def table(self, num_spot_nodes=2049):
    """LocalVolSurface with rows at times_grid[1:] (row k = the function at
    time times_grid[k+1]) on a uniform ln S axis over spot_grid.  The
    splines are flat outside spot_grid, so the function is constant in S
    there but jumps at the ends (its S-derivatives drop to 0): the nodes
    take the one-sided inside values and the edge values are kept apart."""
    lo, hi = float(self._sg.min()), float(self._sg.max())
    x = np.linspace(np.log(lo), np.log(hi), int(num_spot_nodes))
    s = np.exp(x)
    s[0], s[-1] = lo, hi * (1 - 1e-13)
    pts = np.concatenate([[lo * (1 - 1e-9)], s, [hi * (1 + 1e-9)]])
    pts = torch.as_tensor(pts.reshape(-1, 1), dtype=self._dtype,
                          device=self._sg.device)
    rows = np.stack([self(t, pts).reshape(-1).detach().cpu().numpy()
                     for t in self._tg[1:]])
    return LocalVolSurface(self._tg[1:].cpu().numpy(), x[0], x[1] - x[0],
                           rows[:, 1:-1], outside=rows[:, [0, -1]])


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
               spot_grid=None, precompute_iv=False, dtype=None, name=None,
               num_spot_nodes=2049):
    del spot_grid
    self._name = name or 'local_volatility_model'
    self._dim = int(dim)
    self._dtype = dtype if isinstance(dtype, torch.dtype) else (
        torch.float64 if dtype is None or np.dtype(dtype) == np.float64
        else torch.float32)
    self._times_grid = None
    self._precompute_iv = bool(precompute_iv)
    if self._precompute_iv:
      if times_grid is None:
        raise ValueError(
            'When `precompute_iv` is True, both `times_grid` and `spot_grid` '
            'must be supplied')
      self._times_grid = np.asarray(utils.to_numpy(times_grid), np.float64)
    self._local_volatility_fn = local_volatility_fn
    self._num_spot_nodes = int(num_spot_nodes)
    self._rf = risk_free_rate
    self._q = dividend_yield
    self._corr = None if corr_matrix is None else np.asarray(
        utils.to_numpy(corr_matrix), np.float64)
    self._chol = (np.eye(self._dim) if self._corr is None
                  else np.linalg.cholesky(self._corr))
    self._cache = None

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

  def _lv(self, t, x):
    """sigma(t, e^x), computed once per (t, x) for the drift and the
    volatility of one step."""
    key = (float(t), x.data_ptr(), x._version, tuple(x.shape))
    if self._cache is not None and self._cache[0] == key:
      return self._cache[1]
    s = self._local_volatility_fn(t, torch.exp(x))
    self._cache = (key, s)
    return s

  def drift_fn(self):
    """drift(t, X) = r - q - sigma(t, e^X)^2 / 2 (ln S coordinates)."""

    def _drift(t, x):
      s = self._lv(t, x)
      return self._rate(self._rf, t, x) - self._rate(self._q, t, x) - s * s / 2
    spec = self._spec()
    return engine.tag(_drift, spec) if spec else _drift

  def volatility_fn(self):
    """vol(t, X) = sigma(t, e^X)[..., None] * chol(corr)."""
    chol = self._chol

    def _vol(t, x):
      s = self._lv(t, x)
      return s.unsqueeze(-1) * torch.as_tensor(chol, dtype=x.dtype,
                                               device=x.device)
    spec = self._spec()
    return engine.tag(_vol, spec) if spec else _vol

  def _spec(self):
    """('local_vol', table) when the fused kernel applies, else None."""
    if not hasattr(self, '_spec_cache'):
      spec = None
      lv = self._local_volatility_fn
      if self._dim == 1 and isinstance(lv, LocalVolSurface):
        r = _const_rate(self._rf, 1)
        q = _const_rate(self._q, 1)
        if r is not None and q is not None:
          spec = ('local_vol', engine.LocalVolTable(
              lv.times, lv.log_spot_min, lv.log_spot_step, lv.vols,
              float(r[0] - q[0]), lv.outside))
      elif (self._dim == 1 and self._precompute_iv and
            isinstance(lv, _PrecomputedDupire)):
        drift = lv.constant_rate()
        if drift is not None:
          tab = lv.table(self._num_spot_nodes)
          spec = ('local_vol', engine.LocalVolTable(
              tab.times, tab.log_spot_min, tab.log_spot_step, tab.vols,
              drift, tab.outside))
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
      x0 = torch.log(utils.device_tensor(initial_state, self._dtype, dev))
    x = euler_sampling.sample(
        self._dim, self.drift_fn(), self.volatility_fn(), times,
        time_step=time_step, num_time_steps=num_time_steps,
        num_samples=num_samples, initial_state=x0, random_type=random_type,
        seed=seed, skip=skip, times_grid=times_grid,
        normal_draws=normal_draws, watch_params=watch_params,
        validate_args=validate_args, dtype=self._dtype)
    return torch.exp(x)

  @classmethod
  def from_market_data(cls, dim, valuation_date, expiry_dates, strikes,
                       implied_volatilities, spot, discount_factor_fn,
                       dividend_yield=None, local_volatility_from_iv=True,
                       times_grid=None, spot_grid=None, precompute_iv=False,
                       dtype=None, name=None, num_spot_nodes=2049):
    """Local volatility from implied volatilities on (expiry_dates [dim,
    n_expiries], strikes [dim, n_expiries, n_strikes]) nodes, interpolated
    in (expiry, log forward moneyness) (local_volatility_model.py:542-696).
    """
    del name
    if precompute_iv and (times_grid is None or spot_grid is None):
      raise ValueError(
          'When `precompute_iv` is True, both `times_grid` and `spot_grid` '
          'must be supplied')
    dtype = _torch_dtype(dtype, spot)
    spot = as_tensor(spot, dtype)
    dividend_yield = as_tensor([0.0] if dividend_yield is None
                               else dividend_yield, dtype)
    rate_fn = _rate_from_discount_factor(discount_factor_fn)
    expiry_times = vs_lib.daycount_actual_365_fixed(valuation_date,
                                                    expiry_dates, dtype)
    strikes = as_tensor(strikes, dtype)
    ivs = as_tensor(implied_volatilities, dtype)

    def _log_forward_moneyness(times, k):
      r = torch.as_tensor(rate_fn(times), dtype=dtype).squeeze()
      return (torch.log(divide_no_nan(k, spot.reshape(dim, 1, 1))) -
              ((r - dividend_yield) * times).unsqueeze(-1))

    interpolator = interpolation_2d.Interpolation2D(
        expiry_times, _log_forward_moneyness(expiry_times, strikes), ivs,
        dtype=dtype)

    def _log_moneyness_2d_interpolator(times, k):
      r = torch.as_tensor(rate_fn(times), dtype=dtype, device=k.device)
      times = torch.as_tensor(times, dtype=dtype, device=k.device)
      m = (torch.log(k / spot.to(k.device)) -
           (r - dividend_yield.to(k.device)) * times)
      mt = m.transpose(0, 1) if m.dim() == 2 else m
      tt = times.expand(mt.shape)
      out = interpolator.interpolate(tt, mt)
      return out.transpose(0, 1) if out.dim() == 2 else out

    surface = vs_lib.VolatilitySurface(
        valuation_date, expiry_dates, strikes, ivs,
        interpolator=_log_moneyness_2d_interpolator, dtype=dtype)
    return cls._from_surface(dim, spot, surface, discount_factor_fn,
                             dividend_yield, rate_fn, local_volatility_from_iv,
                             times_grid, spot_grid, precompute_iv, dtype,
                             num_spot_nodes)

  @classmethod
  def from_volatility_surface(cls, dim, spot, implied_volatility_surface,
                              discount_factor_fn, dividend_yield=None,
                              local_volatility_from_iv=True,
                              times_grid=None, spot_grid=None,
                              precompute_iv=False, dtype=None, name=None,
                              num_spot_nodes=2049):
    """Local volatility from an object with `volatility(strike,
    expiry_times)` (local_volatility_model.py:698-800)."""
    del name
    if precompute_iv and (times_grid is None or spot_grid is None):
      raise ValueError(
          'When `precompute_iv` is True, both `times_grid` and `spot_grid` '
          'must be supplied')
    dtype = _torch_dtype(dtype, spot)
    dividend_yield = as_tensor([0.0] if dividend_yield is None
                               else dividend_yield, dtype)
    rate_fn = _rate_from_discount_factor(discount_factor_fn)
    return cls._from_surface(dim, as_tensor(spot, dtype),
                             implied_volatility_surface, discount_factor_fn,
                             dividend_yield, rate_fn, local_volatility_from_iv,
                             times_grid, spot_grid, precompute_iv, dtype,
                             num_spot_nodes)

  @classmethod
  def _from_surface(cls, dim, spot, surface, discount_factor_fn,
                    dividend_yield, rate_fn, from_iv, times_grid, spot_grid,
                    precompute_iv, dtype, num_spot_nodes):
    if precompute_iv:
      lv = _PrecomputedDupire(spot, surface.volatility, discount_factor_fn,
                              dividend_yield, times_grid, spot_grid, dim,
                              dtype)
    else:
      lv = functools.partial(
          _dupire_local_volatility_iv if from_iv
          else _dupire_local_volatility_prices,
          initial_spot_price=spot, discount_factor_fn=discount_factor_fn,
          dividend_yield=dividend_yield,
          implied_volatility_surface=surface.volatility)
    return cls(dim, risk_free_rate=rate_fn, dividend_yield=dividend_yield,
               local_volatility_fn=lv, times_grid=times_grid,
               spot_grid=spot_grid, precompute_iv=precompute_iv, dtype=dtype,
               num_spot_nodes=num_spot_nodes)


def _torch_dtype(dtype, like):
  if isinstance(dtype, torch.dtype):
    return dtype
  if dtype is not None:
    return torch.float32 if np.dtype(dtype) == np.float32 else torch.float64
  if isinstance(like, torch.Tensor):
    return like.dtype
  return torch.float64
