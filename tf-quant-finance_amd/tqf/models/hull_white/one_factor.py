"""One-factor Hull-White model (mirror of tff.models.hull_white.
HullWhiteModel1F / VectorHullWhiteModel with dim = 1).

  HullWhiteModel1F              tff/models/hull_white/one_factor.py:26-137
  sample_paths                  vector_hull_white.py:319-449, 641-781
  sample_discount_curve_paths   :451-592
  discount_bond_price           :594-635
  bond reconstitution           :783-814
  exact-scheme integrals        :816-955
  grid                          :996-1041

dr = (theta(t) - a(t) r) dt + sigma(t) dW with theta fitted to the initial
curve P(0, t) = exp(-R(t) t).  With a constant mean reversion and a
piecewise-constant volatility the deviation x = r - f(0, t) is Gaussian
and steps exactly:
    x' = e^{-a dt} x + m + s z,
m and s^2 the conditional mean and variance of the reference's integrals
(Brigo-Mercurio 2007, 3.3.1).  Those per-step constants are computed once
on the host in float64 and the path loop runs in the fused HIP kernel
(tqf_euler with TQF_MODEL_HULL_WHITE: x in a register, one FMA pair per
step, rates recorded as x + f(0, t)).  A mean reversion with jumps or any
non-piecewise-constant parameter takes the reference's Euler route (drift
f'(0,t) + a f(0,t) + sigma^2 (1 - e^{-2at}) / (2a) - a r) through
`euler_sampling.sample`.
"""
import numpy as np
import torch

from tqf import _native
from tqf.engine import euler as engine
from tqf.math.piecewise import PiecewiseConstantFunc
from tqf.math.random.multivariate_normal import RandomType
from tqf.models import euler_sampling
from tqf.models import utils

__all__ = ['HullWhiteModel1F']


def _input_type(param, name):
  """(param, sample_with_generic, is_piecewise_constant)
  (vector_hull_white.py:1044-1088)."""
  if getattr(param, 'is_piecewise_constant', False):
    jumps = np.asarray(param.jump_locations()).reshape(-1)
    generic = name == 'mean_reversion' and jumps.shape[0] > 0
    return param, generic, True
  if callable(param):
    return param, True, False
  v = np.asarray(utils.to_numpy(param), np.float64).reshape(-1)
  if v.shape[0] != 1:
    raise ValueError('Length of {} ({}) should be the same as `dims`(1).'
                     .format(name, v.shape[0]))
  return PiecewiseConstantFunc([], v), False, True


def _pw(param, t):
  """Values of a piecewise-constant parameter at times t (numpy)."""
  return np.asarray(param(np.asarray(t, np.float64)), np.float64)


def _y_int(t0, t, vol, k):
  """int_t0^t sigma^2 e^{2ks} ds on one constant piece."""
  return vol * vol / (2 * k) * (np.exp(2 * k * t) - np.exp(2 * k * t0))


def _ex_int(t0, t, vol, k, y_t0):
  """int_t0^t e^{ks} y(s) ds on one constant piece, with
  y(s) = e^{-2ks} (y_t0 + int_t0^s sigma^2 e^{2ku} du)."""
  value = (np.exp(k * t) - np.exp(k * t0) + np.exp(2 * k * t0) *
           (np.exp(-k * t) - np.exp(-k * t0)))
  return value * vol**2 / (2 * k * k) + y_t0 * (np.exp(-k * t0) -
                                                 np.exp(-k * t)) / k


class HullWhiteModel1F:
  """r(t) with mean reversion a(t), volatility sigma(t) and the initial
  zero rates R(t) of `initial_discount_rate_fn` (P(0,t) = e^{-R(t) t})."""

  def __init__(self, mean_reversion, volatility, initial_discount_rate_fn,
               dtype=None, name=None):
    self._name = name or 'hull_white_one_factor'
    if dtype is None:
      self._dtype = torch.float32
    elif isinstance(dtype, torch.dtype):
      self._dtype = dtype
    else:
      self._dtype = (torch.float64 if np.dtype(dtype) == np.float64
                     else torch.float32)
    self._rate_fn = initial_discount_rate_fn
    self._mean_reversion, g1, p1 = _input_type(mean_reversion,
                                               'mean_reversion')
    self._volatility, g2, p2 = _input_type(volatility, 'volatility')
    self._sample_with_generic = g1 or g2
    self._is_piecewise_constant = p1 and p2
    if self._is_piecewise_constant:
      self._exact_setup()

  # ---------------------------------------------------------------- curve

  def _zero_rate(self, t):
    """R(t) on the host (float64, shape of t)."""
    t = torch.as_tensor(np.array(t, np.float64))
    r = torch.as_tensor(self._rate_fn(t), dtype=torch.float64)
    return np.broadcast_to(r.detach().cpu().numpy().reshape(
        r.shape[:t.dim()] if r.dim() > t.dim() else r.shape),
        t.shape).astype(np.float64)

  def _fwd0(self, t):
    """f(0, t) = d/dt (R(t) t) (vector_hull_white.py:209-225)."""
    t = torch.as_tensor(np.array(t, np.float64))  # a copy: broadcast views
    # alias memory, which forward-mode AD rejects

    def log_bond(x):
      r = torch.as_tensor(self._rate_fn(x), dtype=torch.float64)
      if r.dim() > x.dim():
        r = r.reshape(x.shape)
      return r * x
    f = torch.func.jvp(log_bond, (t,), (torch.ones_like(t),))[1]
    return np.broadcast_to(f.detach().cpu().numpy(), t.shape).astype(
        np.float64)

  def instant_forward_rate(self, t):
    """f(0, t) as a tensor of the model's dtype, shape t.shape + [1]."""
    return torch.as_tensor(self._fwd0(utils.to_numpy(t)),
                           dtype=self._dtype)[..., None]

  def mean_reversion(self):
    return self._mean_reversion

  def volatility(self):
    return self._volatility

  def dtype(self):
    return self._dtype

  def name(self):
    return self._name

  def dim(self):
    return 1

  # ------------------------------------------------------- exact scheme

  def _exact_setup(self):
    """Knot-wise integrals over the parameters' jump locations
    (vector_hull_white.py:816-855)."""
    jumps = np.sort(np.concatenate([
        np.asarray(self._volatility.jump_locations(), np.float64).reshape(-1),
        np.asarray(self._mean_reversion.jump_locations(),
                   np.float64).reshape(-1)]))
    self._jumps = jumps
    jv_vol = _pw(self._volatility, jumps)
    jv_mr = _pw(self._mean_reversion, jumps)
    padded = np.concatenate([[0.0], jumps[:-1]])
    y_between = _y_int(padded, jumps, jv_vol, jv_mr)
    self._y_at = np.concatenate([[0.0], np.cumsum(y_between)])
    ex_between = _ex_int(padded, jumps, jv_vol, jv_mr, self._y_at[:-1])
    self._ex_at = np.concatenate([[0.0], np.cumsum(ex_between)])
    self._vn = np.concatenate([[0.0], jumps])

  def _knot(self, t):
    t = np.asarray(t, np.float64)
    return t, np.searchsorted(self._jumps, t, side='left'), _pw(
        self._mean_reversion, t), _pw(self._volatility, t)

  def _compute_yt(self, t):
    """y(t) = e^{-2at} int_0^t sigma^2 e^{2as} ds (:857-874)."""
    t, idx, mr, vol = self._knot(t)
    y = _y_int(self._vn[idx], t, vol, mr) + self._y_at[idx]
    return np.exp(-2 * mr * t) * y

  def _step_constants(self, t):
    """(decay, mean, sd) of the exact steps between consecutive times
    (:876-955 and the update of :738-760)."""
    t, idx, mr, vol = self._knot(t)
    e_t = _ex_int(self._vn[idx], t, vol, mr, self._y_at[idx]) + \
        self._ex_at[idx]
    mean = (e_t[1:] - e_t[:-1]) * np.exp(-mr[1:] * t[1:])
    v_t = _y_int(self._vn[idx], t, vol, mr) + self._y_at[idx]
    var = (v_t[1:] - v_t[:-1]) * np.exp(-2 * mr[1:] * t[1:])
    decay = np.exp(-mr[1:] * (t[1:] - t[:-1]))
    sd = np.sqrt(np.maximum(var, 0.0))
    return decay, mean, sd

  def _kernel_constants(self, t):
    decay, mean, sd = self._step_constants(t)
    f = self._fwd0(t)
    return decay, mean, sd, f[1:], f[0]

  def _prepare_grid(self, times, times_grid):
    """All step times and the record mask (:996-1041): times, 0 and the
    parameters' jumps, sorted (duplicates kept), or times_grid with each
    requested time moved to its nearest grid point."""
    times = np.asarray(times, np.float64)
    if times_grid is None:
      extra = [self._jumps] if self._is_piecewise_constant else []
      all_times = np.sort(np.concatenate([[0.0], times] + extra))
      idx = np.searchsorted(all_times, times, side='left')
    else:
      all_times = np.asarray(utils.to_numpy(times_grid), np.float64)
      idx = np.searchsorted(all_times, times, side='left')
      up = np.minimum(idx, all_times.shape[0] - 1)
      down = np.maximum(idx - 1, 0)
      idx = np.where(np.abs(all_times[down] - times) >
                     np.abs(all_times[up] - times), up, down)
    mask = np.zeros(all_times.shape[0], bool)
    mask[idx] = True
    return all_times, mask

  # ------------------------------------------------------------ sampling

  def sample_paths(self, times, num_samples=1, random_type=None, seed=None,
                   skip=0, time_step=None, times_grid=None,
                   normal_draws=None, validate_args=False, name=None):
    """Short-rate paths [num_samples, len(times), 1]."""
    del name
    times_np = np.asarray(utils.to_numpy(times), np.float64)
    if times_np.ndim != 1:
      raise ValueError('`times` should be a rank 1 Tensor. '
                       'Rank is {} instead.'.format(times_np.ndim))
    if self._sample_with_generic:
      if time_step is None and times_grid is None:
        raise ValueError(
            'Either `time_step` or `times_grid` has to be specified when '
            'at least one of the parameters is a generic callable.')
      return euler_sampling.sample(
          dim=1, drift_fn=self._drift_fn, volatility_fn=self._volatility_fn,
          times=times_np, time_step=time_step, num_samples=num_samples,
          initial_state=[float(self._fwd0(0.0))], random_type=random_type,
          seed=seed, skip=skip, times_grid=times_grid,
          normal_draws=normal_draws, dtype=self._dtype)
    dev = _native.device()
    all_times, mask = self._prepare_grid(times_np, times_grid)
    steps = all_times.shape[0] - 1
    draws = None
    if normal_draws is not None:
      draws = utils.device_tensor(normal_draws, self._dtype, dev)
      if draws.shape[-1] != 1:
        raise ValueError(
            '`dim` should be equal to `normal_draws.shape[2]` but are '
            '1 and {0} respectively'.format(draws.shape[-1]))
      if validate_args and draws.shape[1] != steps:
        raise ValueError(
            '`tf.shape(normal_draws)[1]` should be equal to the number of '
            'all `times` plus the number of all jumps of the piecewise '
            'constant parameters.')
      num_samples = draws.shape[0]
    elif random_type in (RandomType.HALTON, RandomType.HALTON_RANDOMIZED):
      draws = utils.generate_mc_normal_draws(
          num_normal_draws=1, num_time_steps=steps,
          num_sample_paths=num_samples, random_type=random_type, seed=seed,
          skip=skip, dtype=self._dtype).permute(1, 0, 2)
    rng = engine.RngSpec.from_random_type(random_type, seed, skip, draws)
    spec = ('hull_white', self._kernel_constants)
    return engine.run(spec, all_times, mask, times_np.shape[0],
                      torch.zeros(1, dtype=self._dtype, device=dev),
                      int(num_samples), rng, self._dtype, mode='paths',
                      dev=dev)

  def _bond_constants(self, times, maturities):
    """Host float64 parts of the bond reconstitution for P(t, T) =
    P(0,T)/P(0,t) exp(-x G - y G^2 / 2), x = r - f(0, t)."""
    t = np.asarray(times, np.float64)
    big_t = np.asarray(maturities, np.float64)
    t_b, big_t_b = np.broadcast_arrays(t, big_t)
    mr = _pw(self._mean_reversion, t_b)
    f0 = self._fwd0(t_b)
    p0t = np.exp(-self._zero_rate(t_b) * t_b)
    p0big = np.exp(-self._zero_rate(big_t_b) * big_t_b) / p0t
    g = (1.0 - np.exp(-mr * (big_t_b - t_b))) / mr
    y = self._compute_yt(t_b.reshape(-1)).reshape(t_b.shape)
    return f0, p0big, g, y

  def _bond_prices(self, short_rate, f0, p0big, g, y):
    dev = short_rate.device
    as_t = lambda a: torch.as_tensor(a, dtype=short_rate.dtype, device=dev)
    x = short_rate - as_t(f0)
    gt = as_t(g)
    return as_t(p0big) * torch.exp(-x * gt - 0.5 * as_t(y) * gt * gt)

  def sample_discount_curve_paths(self, times, curve_times, num_samples=1,
                                  random_type=None, seed=None, skip=0,
                                  time_step=None, times_grid=None,
                                  normal_draws=None, validate_args=False,
                                  name=None):
    """(P(t_k, t_k + tau_m) [num_samples, m, k, 1], r [num_samples, k, 1])
    (vector_hull_white.py:451-592)."""
    del name
    if not self._is_piecewise_constant:
      raise ValueError('All paramaters `mean_reversion`, `volatility`, and '
                       '`corr_matrix`must be piecewise constant functions.')
    times_np = np.asarray(utils.to_numpy(times), np.float64).reshape(-1)
    curve = np.asarray(utils.to_numpy(curve_times), np.float64).reshape(-1)
    rate_paths = self.sample_paths(
        times_np, num_samples=num_samples, random_type=random_type,
        seed=seed, skip=skip, time_step=time_step, times_grid=times_grid,
        normal_draws=normal_draws, validate_args=validate_args)
    t = times_np[None, :]
    f0, p0big, g, y = self._bond_constants(t, t + curve[:, None])  # [m, k]
    p = self._bond_prices(rate_paths[:, None, :, 0], f0, p0big, g, y)
    return p.unsqueeze(-1), rate_paths

  def discount_bond_price(self, short_rate, times, maturities, name=None):
    """P(t, T) given r(t): short_rate batch + [1], times and maturities
    batch (vector_hull_white.py:594-635)."""
    del name
    dev = _native.device()
    r = utils.device_tensor(short_rate, self._dtype, dev)
    t = np.asarray(utils.to_numpy(times), np.float64)
    big_t = np.asarray(utils.to_numpy(maturities), np.float64)
    f0, p0big, g, y = self._bond_constants(t, big_t)
    return self._bond_prices(r, f0[..., None], p0big[..., None],
                             g[..., None], y[..., None])

  # ----------------------------------------------------- Euler (generic)

  def _param_at(self, param, t):
    if getattr(param, 'is_piecewise_constant', False):
      return float(_pw(param, np.array([t]))[0])
    return float(utils.to_numpy(param(torch.tensor(t, dtype=torch.float64))))

  def _fwd0_and_slope(self, t):
    """f(0, t) and df(0, t)/dt by nested forward-mode AD of R(t) t."""
    tc = torch.tensor(float(t), dtype=torch.float64)

    def log_bond(x):
      r = torch.as_tensor(self._rate_fn(x), dtype=torch.float64)
      return r.reshape(x.shape) * x

    def fwd(x):
      return torch.func.jvp(log_bond, (x,), (torch.ones_like(x),))[1]
    f, fp = torch.func.jvp(fwd, (tc,), (torch.ones_like(tc),))
    return float(f), float(fp)

  def _drift_fn(self, t, x):
    """f'(0,t) + a f(0,t) + sigma^2 (1 - e^{-2at}) / (2a) - a x
    (vector_hull_white.py:292-308)."""
    tt = float(t)
    a = self._param_at(self._mean_reversion, tt)
    s = self._param_at(self._volatility, tt)
    f, fp = self._fwd0_and_slope(tt)
    c = fp + a * f + s * s / 2 / a * (1 - np.exp(-2 * a * tt))
    return c - a * x

  def _volatility_fn(self, t, x):
    s = self._param_at(self._volatility, float(t))
    return torch.full(x.shape[:-1] + (1, 1), s, dtype=x.dtype,
                      device=x.device)
