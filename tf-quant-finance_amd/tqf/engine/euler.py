"""Launch layer of the fused Euler-Maruyama kernels (tqf_euler).

A *coefficient spec* names a model whose drift/volatility the kernels know:
  ('gbm', mu, sigma)                 dX = mu X dt + sigma X dW
  ('const', a, b)                    dX = a dt + b dW
  ('heston', r, kappa, theta, xi, rho, variant)   (log S, v)
  ('local_vol', LocalVolTable)       d ln S = (r - q - s^2/2) dt + s dW
  ('hull_white', coef_fn)            r = x + f(0, t), x' = e x + m + s z
  ('heston_qe', coef[steps, 12])     Andersen QE with per-step constants
Callables returned by the tqf model classes carry their spec in the
`_tqf_spec` attribute, so `euler_sampling.sample` can route them here.
"""
import ctypes

import numpy as np
import torch

from tqf import _native
from tqf.math.random import sobol as sobol_lib
from tqf.math.random.multivariate_normal import RandomType, seed_to_key
from tqf.models import utils

__all__ = ['tag', 'spec_of', 'model_dim', 'run', 'prepare', 'Launch',
           'RngSpec', 'LocalVolTable']


def tag(fn, spec):
  fn._tqf_spec = spec
  return fn


def spec_of(drift_fn, volatility_fn):
  a = getattr(drift_fn, '_tqf_spec', None)
  b = getattr(volatility_fn, '_tqf_spec', None)
  if a is None or b is None:
    return None
  if a[0] != b[0]:
    return None
  if a[0] == 'local_vol':
    return a if a[1] is b[1] else None
  return a if tuple(a) == tuple(b) else None


def model_dim(spec):
  return 2 if spec[0] in ('heston', 'heston_qe') else 1


class LocalVolTable:
  """sigma(t, S) on a (t_i, ln S_j) grid with a uniform log-spot axis.

  Lookups are bilinear in (t, ln S).  Outside the nodes the value is flat,
  or, with `outside` ([nt, 2]: below the first node, above the last), those
  values are used as they are (a surface that jumps at the ends of its
  spot range).  The kernel stages the table in LDS when it fits.
  """

  def __init__(self, times, log_spot_min, log_spot_step, vols, drift_rate,
               outside=None):
    self.times = np.asarray(times, np.float64)
    self.x0 = float(log_spot_min)
    self.dx = float(log_spot_step)
    self.vols = np.asarray(vols, np.float64)
    self.drift_rate = float(drift_rate)
    self.outside = None if outside is None else np.asarray(outside,
                                                           np.float64)
    if self.vols.shape != (self.times.shape[0], self.vols.shape[1]):
      raise ValueError('vols must have shape [len(times), n_spots]')
    if self.outside is not None and self.outside.shape != (
        self.times.shape[0], 2):
      raise ValueError('outside must have shape [len(times), 2]')

  def device_rows(self):
    """The kernel's row layout: [nt][nx] or, with edge values,
    [nt][1 + nx + 1]."""
    if self.outside is None:
      return self.vols
    return np.concatenate([self.outside[:, :1], self.vols,
                           self.outside[:, 1:]], 1)

  def sigma(self, t, x):
    """Host evaluation of the table (same interpolation as the kernel)."""
    nt, nx = self.vols.shape
    x = np.asarray(x, np.float64)
    fraw = (x - self.x0) / self.dx
    fx = np.clip(fraw, 0, nx - 1)
    j = np.minimum(fx.astype(np.int64), nx - 2)
    wx = fx - j
    row = max(int(np.searchsorted(self.times, t, side='right')) - 1, 0)
    r1 = min(row + 1, nt - 1)
    wt = 0.0
    if row + 1 < nt and t > self.times[row]:
      wt = (t - self.times[row]) / (self.times[row + 1] - self.times[row])
    s0 = self.vols[row, j] + wx * (self.vols[row, j + 1] - self.vols[row, j])
    s1 = self.vols[r1, j] + wx * (self.vols[r1, j + 1] - self.vols[r1, j])
    if self.outside is not None:
      lo, hi = fraw < 0, fraw > nx - 1
      o = self.outside
      s0 = np.where(lo, o[row, 0], np.where(hi, o[row, 1], s0))
      s1 = np.where(lo, o[r1, 0], np.where(hi, o[r1, 1], s1))
    return s0 + wt * (s1 - s0)


class RngSpec:
  """Where the normals come from.

  kind: 'draws' (tensor [N, steps, d]), 'philox', 'philox_antithetic',
  'sobol' (reference transform, index skip+1+p), 'sobol_shifted' (32-digit
  points from index `first_index`, XOR `shift`).
  """

  def __init__(self, kind, seed=0, stream_id=0, draws=None, skip=0,
               shift=None, first_index=0):
    self.kind = kind
    self.seed = seed
    self.stream_id = stream_id
    self.draws = draws
    self.skip = skip
    self.shift = shift
    self.first_index = first_index

  @classmethod
  def from_random_type(cls, random_type, seed, skip, normal_draws=None):
    if normal_draws is not None:
      return cls('draws', draws=normal_draws)
    if random_type is None:
      random_type = RandomType.PSEUDO
    if random_type in (RandomType.PSEUDO, RandomType.STATELESS):
      return cls('philox', seed=seed_to_key(seed, random_type))
    if random_type in (RandomType.PSEUDO_ANTITHETIC,
                       RandomType.STATELESS_ANTITHETIC):
      return cls('philox_antithetic', seed=seed_to_key(seed, random_type))
    if random_type == RandomType.SOBOL:
      return cls('sobol', skip=int(skip or 0))
    raise NotImplementedError(
        'random type {} is not supported by the engine'.format(random_type))


_RNG_CODES = {'draws': _native.RNG_DRAWS, 'philox': _native.RNG_PHILOX,
              'philox_antithetic': _native.RNG_PHILOX_ANTITHETIC,
              'sobol': _native.RNG_SOBOL,
              'sobol_shifted': _native.RNG_SOBOL_SHIFTED}


_NP_OF = {torch.float32: np.float32, torch.float64: np.float64,
          torch.int32: np.int32}


class Staging:
  """The small per-call host arrays of a launch (grid, step coefficients,
  shift, strikes, discounts) uploaded together: one pinned staging buffer
  per device and stream (rewritten once its previous upload has completed)
  and one asynchronous copy into one device buffer, each array a view of
  it.  `add` records an array and the struct fields that take its device
  address; `flush` uploads and sets them."""

  _host = {}

  def __init__(self, dev):
    self.dev = dev
    self.items = []

  def add(self, arr, tdtype, patches=()):
    a = np.ascontiguousarray(np.asarray(arr).astype(_NP_OF[tdtype],
                                                    copy=False))
    self.items.append((a, tdtype, list(patches)))

  def flush(self, keep):
    if not self.items:
      return []
    offs, total = [], 0
    for a, _, _ in self.items:
      total = (total + 255) & ~255
      offs.append(total)
      total += a.nbytes
    total = max(total, 1)
    stream = torch.cuda.current_stream(self.dev)
    key = (str(self.dev), stream.cuda_stream)
    ent = Staging._host.get(key)
    if ent is None or ent['buf'].numel() < total:
      cap = 1 << max(16, (total - 1).bit_length())
      ent = {'buf': torch.empty(cap, dtype=torch.uint8, pin_memory=True),
             'done': None}
      Staging._host[key] = ent
    if ent['done'] is not None:
      ent['done'].synchronize()
    hb = ent['buf'].numpy()
    for (a, _, _), o in zip(self.items, offs):
      hb[o:o + a.nbytes] = a.reshape(-1).view(np.uint8)
    dbuf = torch.empty(total, dtype=torch.uint8, device=self.dev)
    dbuf.copy_(ent['buf'][:total], non_blocking=True)
    ent['done'] = torch.cuda.Event()
    ent['done'].record(stream)
    keep.append(dbuf)
    outs = []
    for (a, td, patches), o in zip(self.items, offs):
      t = dbuf[o:o + a.nbytes].view(td)
      outs.append(t)
      for obj, field in patches:
        setattr(obj, field, t.data_ptr())
    self.items = []
    return outs


def _model_struct(spec, dtype, dev, keep, step_times=None, step_slots=None,
                  stage=None):
  """The model descriptor; its device arrays go through `stage` when one is
  given (their addresses are set by stage.flush), else are uploaded here."""
  def upload(arr, tdtype, field):
    if stage is not None:
      stage.add(arr, tdtype, [(m, field)])
      return
    t = torch.as_tensor(np.ascontiguousarray(arr), device=dev).to(tdtype)
    keep.append(t)
    setattr(m, field, t.data_ptr())

  m = _native.Model()
  d = model_dim(spec)
  m.dim = d
  if spec[0] == 'gbm':
    m.kind = _native.MODEL_GBM
    m.p[0], m.p[1] = float(spec[1]), float(spec[2])
  elif spec[0] == 'const':
    m.kind = _native.MODEL_CONST
    m.p[0], m.p[1] = float(spec[1]), float(spec[2])
  elif spec[0] == 'heston':
    m.kind = _native.MODEL_HESTON_FT
    for i in range(6):
      m.p[i] = float(spec[1 + i])
    if step_times is not None:
      # per-step rows {-dt/2, kappa theta dt, -kappa dt, r t_{i+1},
      # sqrt(dt), slot bits, 0, 0}
      r, kappa, theta = float(spec[1]), float(spec[2]), float(spec[3])
      npd = utils.np_dtype(dtype)
      t = np.asarray(step_times, np.float64)
      dt = (t[1:] - t[:-1]).astype(npd)
      sdt = np.sqrt(dt).astype(npd)
      dt = dt.astype(np.float64)
      n = dt.shape[0]
      coef = np.zeros((max(n, 1), 8), npd)
      coef[:n, 0] = -0.5 * dt
      coef[:n, 1] = kappa * theta * dt
      coef[:n, 2] = -kappa * dt
      coef[:n, 3] = r * t[1:]
      coef[:n, 4] = sdt
      ib = np.int32 if npd == np.float32 else np.int64
      bits = coef[:, 5].view(ib)
      bits[:] = -1
      bits[:n] = np.asarray(step_slots[:n], ib)
      upload(coef, dtype, 'step_coef')
  elif spec[0] == 'heston_qe':
    m.kind = _native.MODEL_HESTON_QE
    upload(utils.to_numpy(spec[1]), dtype, 'step_coef')
  elif spec[0] == 'hull_white':
    # spec[1](step_times) -> (decay, mean, sd, f(0, t_{i+1}), f(0, t_0)) of
    # the exact steps between consecutive grid times
    m.kind = _native.MODEL_HULL_WHITE
    npd = utils.np_dtype(dtype)
    t = np.asarray(step_times, np.float64)
    decay, mean, sd, f_next, f_first = spec[1](t)
    m.p[0] = float(f_first)
    n = t.shape[0] - 1
    coef = np.zeros((max(n, 1), 8), npd)
    coef[:n, 0] = decay[:n]
    coef[:n, 1] = mean[:n]
    coef[:n, 2] = sd[:n]
    coef[:n, 3] = f_next[:n]
    coef[:n, 4] = 1.0
    ib = np.int32 if npd == np.float32 else np.int64
    bits = coef[:, 5].view(ib)
    bits[:] = -1
    bits[:n] = np.asarray(step_slots[:n], ib)
    upload(coef, dtype, 'step_coef')
  elif spec[0] == 'local_vol':
    tab = spec[1]
    m.kind = _native.MODEL_LOCAL_VOL
    m.p[0] = tab.drift_rate
    m.p[1] = 0.0 if tab.outside is None else 1.0
    rows = tab.device_rows()
    upload(rows, dtype, 'lv_vol')
    upload(utils.to_numpy(tab.times), dtype, 'lv_t')
    m.lv_x0, m.lv_dx = tab.x0, tab.dx
    m.lv_nt, m.lv_nx = rows.shape
  else:
    raise NotImplementedError('unknown model spec %r' % (spec[0],))
  return m


class Launch:
  """A fully prepared tqf_euler call: descriptors and device buffers are
  built once; `launch()` only enqueues the kernels on the current stream."""

  def __init__(self, m, g, r, o, num_samples, x0_t, x0_stride, dtype, dev,
               keep, result):
    self.m, self.g, self.r, self.o = m, g, r, o
    self.num_samples = num_samples
    self.x0_t, self.x0_stride = x0_t, x0_stride
    self.dtype, self.dev = dtype, dev
    self.keep = keep
    self.result = result

  def launch(self):
    _native.call('tqf_euler', ctypes.byref(self.m), ctypes.byref(self.g),
                 ctypes.byref(self.r), ctypes.byref(self.o),
                 ctypes.c_int64(self.num_samples),
                 ctypes.c_void_p(self.x0_t.data_ptr()),
                 ctypes.c_int64(self.x0_stride),
                 ctypes.c_int32(_native.dtype_code(self.dtype)),
                 _native.stream_ptr(self.dev))
    return self.result

  def launch_split(self):
    """Payoff mode in two calls: the path kernel with per-block sums
    (TQF_OUT_PAYOFF_PARTIALS), then tqf_payoff_finalize.  Same result as
    `launch()`; returns the two callables so a caller can time or order
    them separately."""
    if self.o.mode != _native.OUT_PAYOFF:
      raise ValueError('launch_split needs payoff mode')
    o2 = _native.Output.from_buffer_copy(self.o)
    o2.mode = _native.OUT_PAYOFF_PARTIALS
    n_strikes = int(self.o.n_strikes)
    n_slots = int(self.g.n_slots)

    def paths():
      _native.call('tqf_euler', ctypes.byref(self.m), ctypes.byref(self.g),
                   ctypes.byref(self.r), ctypes.byref(o2),
                   ctypes.c_int64(self.num_samples),
                   ctypes.c_void_p(self.x0_t.data_ptr()),
                   ctypes.c_int64(self.x0_stride),
                   ctypes.c_int32(_native.dtype_code(self.dtype)),
                   _native.stream_ptr(self.dev))

    def finalize():
      _native.call('tqf_payoff_finalize', ctypes.c_void_p(self.o.workspace),
                   ctypes.c_int64(self.num_samples), ctypes.c_int32(n_slots),
                   ctypes.c_int32(n_strikes),
                   ctypes.c_void_p(self.o.discount),
                   ctypes.c_void_p(self.o.prices),
                   ctypes.c_void_p(self.o.stderrs),
                   ctypes.c_void_p(self.o.sums), _native.stream_ptr(self.dev))
      return self.result
    return paths, finalize


def run(spec, all_times, keep_mask, num_requested, x0, num_samples, rng,
        dtype, mode='paths', strikes=None, discount=None,
        initial_written=None, dev=None):
  """Runs the fused kernel once (see `prepare`)."""
  return prepare(spec, all_times, keep_mask, num_requested, x0, num_samples,
                 rng, dtype, mode, strikes, discount, initial_written,
                 dev).launch()


def prepare(spec, all_times, keep_mask, num_requested, x0, num_samples, rng,
            dtype, mode='paths', strikes=None, discount=None,
            initial_written=None, dev=None):
  """Builds a Launch of the fused kernel.

  mode 'paths'  -> [N, k, d] (reference layout, contiguous)
  mode 'store'  -> [k, d, N] (coalesced)
  mode 'view'   -> [N, k, d] view of the coalesced [k, d, N] store (the
                   reference's shape and values, written at the path-store
                   rate; .contiguous() gives the reference layout)
  mode 'payoff' -> dict(prices [k, n_strikes], stderrs, sums [k, K, 2])
  """
  dev = dev or _native.device()
  npd = utils.np_dtype(dtype)
  d = model_dim(spec)
  all_times = np.asarray(all_times, npd)
  dt = (all_times[1:] - all_times[:-1]).astype(npd)
  sqrt_dt = np.sqrt(dt).astype(npd)
  slots, init_slot, executed, filled = utils.record_plan(
      keep_mask, num_requested, initial_written)
  n_rng_steps = dt.shape[0]
  keep = []

  # the small arrays go up in one staged copy (addresses set by flush)
  stage = Staging(dev)

  g = _native.Grid()
  stage.add(dt[:max(executed, 1)], dtype, [(g, 'dt')])
  stage.add(sqrt_dt[:max(executed, 1)], dtype, [(g, 'sqrt_dt')])
  stage.add(all_times[1:][:max(executed, 1)], dtype, [(g, 't_next')])
  stage.add(np.concatenate([slots, [-1]]).astype(np.int32), torch.int32,
            [(g, 'slot')])
  g.n_steps = executed
  g.n_slots = int(num_requested)
  g.init_slot = init_slot
  g.n_rng_steps = max(n_rng_steps, executed)

  m = _model_struct(spec, dtype, dev, keep,
                    step_times=all_times[:executed + 1], step_slots=slots,
                    stage=stage)

  r = _native.Rng()
  r.kind = _RNG_CODES[rng.kind]
  r.stream_id = int(rng.stream_id)
  r.seed = int(rng.seed) & 0xFFFFFFFFFFFFFFFF
  if rng.kind == 'draws':
    draws = rng.draws
    if draws.dtype != dtype or draws.device != dev:
      draws = draws.to(device=dev, dtype=dtype)
    if draws.dim() != 3 or draws.shape[2] != d:
      raise ValueError('normal_draws must have shape [num_samples, '
                       'num_time_steps, {}]'.format(d))
    if draws.shape[1] < executed:
      raise ValueError('normal_draws has {} steps, the grid needs {}'.format(
          draws.shape[1], executed))
    keep.append(draws)
    r.draws = draws.data_ptr()
    r.stride_path, r.stride_step = draws.stride(0), draws.stride(1)
    if draws.stride(2) != 1:
      draws = draws.contiguous()
      keep.append(draws)
      r.draws = draws.data_ptr()
      r.stride_path, r.stride_step = draws.stride(0), draws.stride(1)
  elif rng.kind in ('sobol', 'sobol_shifted'):
    ndims = max(n_rng_steps, 1) * d
    if rng.kind == 'sobol':
      nd = sobol_lib.num_digits_for(rng.skip, num_samples)
      first = rng.skip + 1
    else:
      nd = 32
      first = rng.first_index
    v = sobol_lib.device_columns(ndims, nd, dev=dev)
    lo = sobol_lib.device_lo_table(ndims, nd, group=d, dev=dev)
    r.sobol_v, r.sobol_lo = v.data_ptr(), lo.data_ptr()
    if rng.shift is not None:
      sh = np.asarray(rng.shift, np.uint64).astype(np.uint32).view(np.int32)
      if sh.size < ndims:
        raise ValueError('digital shift has %d entries, need %d' %
                         (sh.size, ndims))
      stage.add(sh, torch.int32, [(r, 'sobol_shift')])
    r.sobol_first_index = first
    r.sobol_nd = nd
    r.sobol_ndims = ndims

  x0_t = None
  if not torch.is_tensor(x0) and np.ndim(x0) <= 1:
    # one state for every path: staged with the other small arrays
    x0_h = np.asarray(utils.to_numpy(x0), npd).reshape(-1)
    if x0_h.shape[0] == 1 and d != 1:
      x0_h = np.broadcast_to(x0_h, (d,))
    x0_stride = 0
    x0_at = len(stage.items)
    stage.add(x0_h, dtype)
  else:
    x0_t = torch.as_tensor(x0, dtype=dtype, device=dev)
    if x0_t.dim() == 0:
      x0_t = x0_t.reshape(1)
    if x0_t.dim() == 1:
      if x0_t.shape[0] != d:
        x0_t = x0_t.expand(d) if x0_t.shape[0] == 1 else x0_t
      x0_stride = 0
    else:
      x0_t = x0_t.expand(num_samples, d)
      x0_stride = d
    x0_t = x0_t.contiguous()
    keep.append(x0_t)

  o = _native.Output()
  result = None
  # slots that no executed step records stay zero; when every slot is
  # written the output needs no zero-fill pass
  alloc = torch.empty if len(filled) == int(num_requested) else torch.zeros
  if mode == 'paths':
    o.mode = _native.OUT_PATH_MAJOR
    result = alloc((num_samples, int(num_requested), d), dtype=dtype,
                   device=dev)
    o.paths = result.data_ptr()
  elif mode in ('store', 'view'):
    o.mode = _native.OUT_STEP_MAJOR
    result = alloc((int(num_requested), d, num_samples), dtype=dtype,
                   device=dev)
    o.paths = result.data_ptr()
    if mode == 'view':
      result = result.permute(2, 0, 1)
  elif mode == 'payoff':
    if len(filled) != int(num_requested):
      raise ValueError('every requested time must be reached for payoffs')
    o.mode = _native.OUT_PAYOFF
    kn = np.asarray(utils.to_numpy(strikes), np.float64).reshape(-1)
    stage.add(kn, torch.float64, [(o, 'strikes')])
    o.n_strikes = kn.size
    disc = np.ones(int(num_requested)) if discount is None else discount
    stage.add(np.asarray(utils.to_numpy(disc), np.float64), torch.float64,
              [(o, 'discount')])
    ws_bytes = _native.lib().tqf_euler_workspace_size(
        num_samples, int(num_requested), kn.size)
    ws = torch.empty(max(ws_bytes // 8, 1), dtype=torch.float64, device=dev)
    keep.append(ws)
    o.workspace = ws.data_ptr()
    prices = torch.empty((int(num_requested), kn.size), dtype=torch.float64,
                         device=dev)
    ses = torch.empty_like(prices)
    sums = torch.empty((int(num_requested), kn.size, 2),
                       dtype=torch.float64, device=dev)
    o.prices, o.stderrs, o.sums = (prices.data_ptr(), ses.data_ptr(),
                                   sums.data_ptr())
    result = {'prices': prices, 'stderrs': ses, 'sums': sums}
  else:
    raise ValueError('unknown mode %r' % mode)

  staged = stage.flush(keep)
  if x0_t is None:
    x0_t = staged[x0_at]
  if isinstance(result, dict):
    result['_keep'] = keep
  else:
    # keep the inputs alive until the stream has consumed them
    result._tqf_keep = keep
  return Launch(m, g, r, o, num_samples, x0_t, x0_stride, dtype, dev, keep,
                result)
