"""Learned Longstaff-Schwartz: an MLP continuation value trained per
exercise date with the tqf_mlp_* kernels, judged against the polynomial
regression of the reference (tff/models/longstaff_schwartz/lsm.py).

Backward induction as least_square_mc (lsm.py:258-436), with the
regression step replaced: at each date the network (6 -> 64 -> 64 -> 1,
SiLU, bf16 MFMA) is fitted by Adam to the discounted continuation cashflows
of the in-the-money calibration paths, then predicts the continuation value
of every path.  The price is the discounted mean over the independent
pricing paths, a lower bound for any exercise policy.  The polynomial LSM
price on the same paths is the judge: the MLP price is accepted only when
it is at least poly - 2 SE; otherwise the polynomial price is reported.
Both prices are means over the same pricing paths, so their errors are
strongly correlated: the gain and its standard error come from the per-path
differences of the two policies' discounted cashflows (a paired estimate),
not from the two prices' own standard errors.
"""
import ctypes

import numpy as np
import torch

from tqf import _native
from tqf.models import longstaff_schwartz as lsm_lib

__all__ = ['MlpModel', 'mlp_lsm_price', 'init_params']

N_IN, N_H = 6, 64
N_PARAMS = N_IN * N_H + N_H + N_H * N_H + N_H + N_H + 1


def init_params(seed=0):
  """He-style initialisation in the kernels' parameter layout."""
  rs = np.random.RandomState(seed)
  w1 = rs.standard_normal((N_IN, N_H)) * np.sqrt(2.0 / N_IN)
  w2 = rs.standard_normal((N_H, N_H)) * np.sqrt(2.0 / N_H)
  w3 = rs.standard_normal(N_H) * np.sqrt(1.0 / N_H)
  return np.concatenate([w1.ravel(), np.zeros(N_H), w2.ravel(), np.zeros(N_H),
                         w3, [0.0]]).astype(np.float32)


class MlpModel:
  """Device parameters + Adam state; forward / train-step wrappers."""

  def __init__(self, params=None, seed=0, n_blocks=None, dev=None):
    self.dev = dev or _native.device()
    p = init_params(seed) if params is None else np.asarray(params, np.float32)
    self.params = torch.as_tensor(p, device=self.dev).contiguous()
    self.m = torch.zeros_like(self.params)
    self.v = torch.zeros_like(self.params)
    self.step = 0
    if n_blocks is None:
      # one workgroup per CU: the gradient kernel runs one wave per SIMD
      # (its registers allow no more), so a second round of workgroups
      # only repeats the per-workgroup set-up and reduction (measured
      # 0.039 ms per 65536-sample step against 0.047 ms with two per CU)
      n_blocks = torch.cuda.get_device_properties(
          self.dev).multi_processor_count
    self.n_blocks = int(n_blocks)
    ws = _native.lib().tqf_mlp_workspace_size(self.n_blocks)
    self._ws = torch.empty(ws // 4 + 1, dtype=torch.float32, device=self.dev)
    self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
    self._graphs = {}

  def reset(self, seed=0, params=None):
    """Fresh parameters (init_params(seed) or the given vector) and Adam
    state, in place: captured graphs stay valid."""
    p = init_params(seed) if params is None else np.asarray(params, np.float32)
    self.params.copy_(torch.as_tensor(p, device=self.dev))
    self.m.zero_()
    self.v.zero_()
    self.step = 0

  def forward(self, x, out=None):
    n = x.shape[0]
    if out is None:
      out = torch.empty(n, dtype=torch.float32, device=self.dev)
    _native.call('tqf_mlp_forward', ctypes.c_void_p(self.params.data_ptr()),
                 ctypes.c_void_p(x.data_ptr()), ctypes.c_int64(n),
                 ctypes.c_void_p(out.data_ptr()), _native.stream_ptr(self.dev))
    return out

  def grad(self, x, t, w, first=0, n=None, inv_norm=None):
    """Gradient only (lr = 0): returns (grad [N_PARAMS], loss)."""
    n = x.shape[0] - first if n is None else n
    inv_norm = 1.0 / n if inv_norm is None else inv_norm
    g = torch.empty(N_PARAMS, dtype=torch.float32, device=self.dev)
    self._call(x, t, w, first, n, inv_norm, 0.0, 1, g)
    return g, self.loss

  def train_step(self, x, t, w, first, n, inv_norm, lr):
    self.step += 1
    self._call(x, t, w, first, n, inv_norm, lr, self.step, None)

  def train_date(self, x, t, w, n_cal, batch, steps, lr, graph=True):
    """One exercise date's training: the Adam state reset, then `steps`
    Adam steps over the windows first = k * batch mod (n_cal - batch + 1)
    of the inputs.  With graph=True the whole sequence (two fill kernels
    and 2 * steps launches) is captured once per (inputs, sizes) as a
    hipGraph and replayed: the dates of one pricing repeat the same
    launches on the same buffers, and a replay takes none of the
    per-launch host work or the gaps between dependent launches."""
    if not graph:
      self._date_steps(x, t, w, n_cal, batch, steps, lr)
      return
    key = (x.data_ptr(), t.data_ptr(), w.data_ptr(), int(n_cal), int(batch),
           int(steps), float(lr))
    g = self._graphs.get(key)
    if g is None:
      g = torch.cuda.CUDAGraph()
      with torch.cuda.graph(g):
        self._date_steps(x, t, w, n_cal, batch, steps, lr)
      self._graphs[key] = g
    g.replay()
    self.step = int(steps)

  def _date_steps(self, x, t, w, n_cal, batch, steps, lr):
    # zero_ is a fill kernel (no memset node in a captured graph)
    self.m.zero_()
    self.v.zero_()
    for k in range(int(steps)):
      first = (k * batch) % max(n_cal - batch + 1, 1)
      self._call(x, t, w, first, batch, 1.0 / batch, lr, k + 1, None)
    self.step = int(steps)

  def _call(self, x, t, w, first, n, inv_norm, lr, step, g):
    _native.call('tqf_mlp_train_step', ctypes.c_void_p(self.params.data_ptr()),
                 ctypes.c_void_p(self.m.data_ptr()),
                 ctypes.c_void_p(self.v.data_ptr()),
                 ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(t.data_ptr()),
                 ctypes.c_void_p(w.data_ptr()), ctypes.c_int64(first),
                 ctypes.c_int64(n), ctypes.c_float(inv_norm),
                 ctypes.c_float(lr), ctypes.c_int32(step),
                 ctypes.c_int32(self.n_blocks),
                 ctypes.c_void_p(self._ws.data_ptr()),
                 ctypes.c_void_p(0 if g is None else g.data_ptr()),
                 ctypes.c_void_p(self.loss.data_ptr()),
                 _native.stream_ptr(self.dev))


def _lsm_struct(paths, payoff, strike, degree=2):
  """paths: float64 [N, T, d] with any strides."""
  n, nt, d = paths.shape
  expo = lsm_lib.exponent_grid(degree, d).astype(np.uint8)
  expo_t = torch.as_tensor(expo, device=paths.device).contiguous()
  l = _native.Lsm()
  l.paths = paths.data_ptr()
  l.stride_path, l.stride_time, l.stride_asset = paths.stride()
  l.n_paths = n
  l.n_times = nt
  l.dim = d
  l.payoff = payoff
  l.degree = degree
  l.n_basis = expo.shape[0]
  l.exponents = expo_t.data_ptr()
  l.strike = float(strike)
  return l, expo_t


# The buffers and the model of the last pricing size, kept so that the
# per-date training graphs captured by MlpModel.train_date (which hold their
# device addresses) are replayed by later calls instead of captured again.
# One entry: a call with another device or path count replaces it.  Calls
# share these buffers, so mlp_lsm_price is not for concurrent use from
# several host threads.
_STATE = {}


def _state(dev, n):
  key = (str(dev), int(n))
  st = _STATE.get(key)
  if st is None:
    _STATE.clear()
    f64 = dict(dtype=torch.float64, device=dev)
    st = {
        'cash': torch.empty(n, **f64), 'vals': torch.empty(n, **f64),
        'exv': torch.empty(n, **f64), 'y': torch.empty(n, **f64),
        'mom': torch.empty(_native.lib().tqf_lsm_moments_size(n) // 8 + 1,
                           **f64),
        'x': torch.empty((n, 8), dtype=torch.float32, device=dev),
        'tgt': torch.empty(n, dtype=torch.float32, device=dev),
        'w': torch.empty(n, dtype=torch.float32, device=dev),
        'cont': torch.empty(n, dtype=torch.float32, device=dev),
        'fin': torch.empty((n + 255) // 256 * 2 + 2, **f64),
        'res': torch.empty(2, **f64),
        'model': MlpModel(dev=dev),
    }
    _STATE[key] = st
  return st


def clear_cache():
  """Drops the cached pricing buffers, model and training graphs."""
  _STATE.clear()


_JUDGE_STREAMS = {}


def _judge_stream(dev):
  """One cached side stream per device (default priority)."""
  key = str(dev)
  if key not in _JUDGE_STREAMS:
    _JUDGE_STREAMS[key] = torch.cuda.Stream(device=dev)
  return _JUDGE_STREAMS[key]


def mlp_lsm_price(paths, exercise_idx, strike, discount_factors,
                  num_calibration, payoff=lsm_lib.PAYOFF_MAX_CALL,
                  steps_per_date=200, batch=65536, lr=1e-3, seed=0,
                  poly_degree=2, judge=True, return_path_values=False,
                  graph=True):
  """Prices with the learned continuation value (and the polynomial judge).

  paths: float64 device tensor [N, T, d], rows [0, num_calibration) train
  the models, the rest price.  discount_factors: [n_ex] (unpadded).
  Returns a dict with mlp / poly prices and standard errors, the paired
  gain mlp - poly and its standard error, the accepted price and the
  per-date final training losses; with return_path_values the per-path
  values of both policies at the first exercise date (float64 [N], before
  the final discount).  graph: each date's training replays one captured
  hipGraph (MlpModel.train_date); the first call of a size pays the
  capture (about 0.1 s), later calls of that size replay it.
  """
  dev = paths.device
  paths = lsm_lib.step_major_f64(paths)
  n = paths.shape[0]
  ex = [int(v) for v in exercise_idx]
  df = np.concatenate([[1.0], np.asarray(discount_factors, np.float64)])
  l, keep = _lsm_struct(paths, payoff, strike, poly_degree)
  s = _native.stream_ptr(dev)
  f64 = dict(dtype=torch.float64, device=dev)
  st = _state(dev, n)
  cash, vals, exv, y = st['cash'], st['vals'], st['exv'], st['y']
  mom, x, tgt, w = st['mom'], st['x'], st['tgt'], st['w']
  cont, fin, res = st['cont'], st['fin'], st['res']
  model = st['model']
  model.reset(seed)
  losses = []
  n_cal = int(num_calibration)
  if judge:
    # the polynomial judge depends on the paths only: it runs on a side
    # stream (default priority) while the network trains on this one
    # (the training kernels hold one wave per SIMD); the two meet at the
    # paired statistic
    main = torch.cuda.current_stream(dev)
    side = _judge_stream(dev)
    side.wait_stream(main)
    with torch.cuda.stream(side):
      poly_values = torch.empty(n, **f64)
      poly, poly_se, _ = lsm_lib._native_lsm(
          paths, ex, (payoff, np.array([float(strike)])), poly_degree,
          np.asarray(df).reshape(-1, 1, 1), n_cal, torch.float64, dev,
          return_details=True, path_values=poly_values)
  _native.call('tqf_lsm_terminal', ctypes.byref(l), ctypes.c_int32(ex[-1]),
               ctypes.c_void_p(cash.data_ptr()),
               ctypes.c_void_p(vals.data_ptr()), s)
  batch = min(int(batch), n_cal)
  for e in range(len(ex) - 1, 0, -1):
    ratio = df[e + 1] / df[e]
    _native.call('tqf_lsm_date_inputs', ctypes.byref(l),
                 ctypes.c_int32(ex[e - 1]), ctypes.c_double(ratio),
                 ctypes.c_void_p(cash.data_ptr()),
                 ctypes.c_void_p(vals.data_ptr()),
                 ctypes.c_void_p(exv.data_ptr()),
                 ctypes.c_void_p(y.data_ptr()),
                 ctypes.c_void_p(mom.data_ptr()),
                 ctypes.c_void_p(x.data_ptr()),
                 ctypes.c_void_p(tgt.data_ptr()),
                 ctypes.c_void_p(w.data_ptr()), s)
    model.train_date(x, tgt, w, n_cal, batch, steps_per_date, lr,
                     graph=graph)
    losses.append(model.loss.clone())
    model.forward(x, cont)
    _native.call('tqf_lsm_update_cont', ctypes.c_int64(n),
                 ctypes.c_double(strike), ctypes.c_double(ratio),
                 ctypes.c_void_p(exv.data_ptr()),
                 ctypes.c_void_p(cont.data_ptr()),
                 ctypes.c_void_p(cash.data_ptr()),
                 ctypes.c_void_p(vals.data_ptr()), s)
  _native.call('tqf_lsm_final', ctypes.c_void_p(cash.data_ptr()),
               ctypes.c_void_p(vals.data_ptr()), ctypes.c_double(df[1] / df[0]),
               ctypes.c_int64(n_cal), ctypes.c_int64(n),
               ctypes.c_void_p(fin.data_ptr()), ctypes.c_void_p(res.data_ptr()),
               s)
  mlp_price, mlp_se = [float(v) for v in res.cpu().numpy()]
  out = {'mlp_price': mlp_price, 'mlp_se': mlp_se,
         'losses': [float(t.item()) for t in losses]}
  if judge:
    main.wait_stream(side)
    # read on this stream from here on: their blocks are not handed back
    # to the side stream before these reads are done
    for t in (poly_values, poly, poly_se):
      t.record_stream(main)
    mlp_values = cash + vals
    poly_price = float(poly[0])
    poly_se = float(poly_se[0])
    # paired: mean and standard error of df0 (mlp_p - poly_p) over the
    # pricing rows, by the same kernel that reduces each price
    diff = mlp_values - poly_values
    _native.call('tqf_lsm_final', ctypes.c_void_p(diff.data_ptr()),
                 ctypes.c_void_p(torch.zeros_like(diff).data_ptr()),
                 ctypes.c_double(df[1] / df[0]), ctypes.c_int64(n_cal),
                 ctypes.c_int64(n), ctypes.c_void_p(fin.data_ptr()),
                 ctypes.c_void_p(res.data_ptr()), s)
    gain, gain_se = [float(v) for v in res.cpu().numpy()]
    accepted = gain >= -2 * gain_se
    out.update({'poly_price': poly_price, 'poly_se': poly_se,
                'accepted': bool(accepted),
                'price': mlp_price if accepted else poly_price,
                'gain': gain, 'gain_se': gain_se,
                'gain_se_unpaired': float(np.hypot(mlp_se, poly_se))})
    if return_path_values:
      out['mlp_path_values'] = mlp_values
      out['poly_path_values'] = poly_values
  return out


def bermudan_max_call_paths(n_paths, dim=5, s0=100.0, sigma=0.2, corr=0.3,
                            r=0.05, q=0.1, maturity=3.0, n_dates=9,
                            steps_per_date=20, seed=17):
  """Correlated GBM paths at t = 0 and the exercise dates, float64
  [n_paths, n_dates + 1, dim] (a view of step-major memory [T][d][N]),
  simulated over n_dates * steps_per_date exact steps by tqf_gbm_exact
  (Philox).  Also returns the discount factors of the exercise dates."""
  from tqf.engine import euler as engine
  from tqf.models.geometric_brownian_motion import run_exact
  steps = n_dates * steps_per_date
  grid = np.linspace(0.0, maturity, steps + 1)
  keep = np.zeros(steps + 1, bool)
  keep[::steps_per_date] = True
  c = np.full((dim, dim), corr) + (1 - corr) * np.eye(dim)
  drift = np.full(dim, (r - q) - 0.5 * sigma * sigma)
  vol = np.full(dim, sigma)
  out = run_exact(grid[1:], drift, vol, np.linalg.cholesky(c),
                  np.full(dim, s0), int(n_paths),
                  engine.RngSpec('philox', seed=seed), torch.float32,
                  keep_mask=keep, mode='store')  # [T][d][N]
  dates = grid[keep][1:]
  return out.to(torch.float64).permute(2, 0, 1), np.exp(-r * dates)
