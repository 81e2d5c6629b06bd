"""American options by Longstaff-Schwartz under Black-Scholes (mirror of the
pricing platform's `bs_lsm_price`, tff/experimental/pricing_platform/
framework/equity_instruments/american_option/utils.py:26-169).

The reference draws one set of unit GBM paths exp(-t/2 + W(t)) on the
normalised exercise grid t = linspace(0, 1, num_exercise_times) and maps
them to every option of the batch as spot-scaled paths
S/S0 = exp(sigma sqrt(T) W(t) + (r - sigma^2/2) t T) (utils.py:103-136),
with r = -ln(df)/T; each option is then priced by least_square_mc with the
degree-2 polynomial basis on ln S and the call/put payoff on S
(utils.py:137-167), and the result is scaled back by the spot.

Here the unit paths come from the exact GBM kernel (tqf_gbm_exact) with
the engine's antithetic Philox stream (the reference's
STATELESS_ANTITHETIC), the per-option mapping to ln(S / S0) is one
elementwise pass on the GPU, and the regression runs in the FP64 LSM
kernels (tqf_lsm_price on log prices: basis on ln S, payoff on S).  The
options' regressions are independent: each runs on its own stream (up to
four, default priority), so one option's single-workgroup solves overlap
the other options' bandwidth-bound date passes, and the prices come back
in one copy.  A user basis_fn falls back to the torch LSM with the same
log-path / payoff split.
"""
import numpy as np
import torch

from tqf import _native
from tqf.engine import euler as engine
from tqf.math.random.multivariate_normal import RandomType
from tqf.models import longstaff_schwartz as lsm_lib
from tqf.models import utils

__all__ = ['bs_lsm_price']


_STREAMS = {}
_MAX_STREAMS = 4


def _streams(dev, n):
  """Up to _MAX_STREAMS side streams of `dev` (default priority), cached."""
  key = str(dev)
  pool = _STREAMS.setdefault(key, [])
  while len(pool) < min(n, _MAX_STREAMS):
    pool.append(torch.cuda.Stream(device=dev))
  return pool[:min(n, _MAX_STREAMS)]


def _vec(v, n, dtype):
  a = np.asarray(utils.to_numpy(v), np.float64).reshape(-1)
  return np.broadcast_to(a, [n]).astype(np.float64)


def bs_lsm_price(spots, expiry_times, strikes, volatility, discount_factors,
                 num_samples=100000, num_exercise_times=100, basis_fn=None,
                 seed=(1, 2), is_call_option=True,
                 num_calibration_samples=None, dtype=None, name=None):
  """American option prices [batch] (same arguments as the reference)."""
  del name
  dtype = dtype or torch.float64
  if not isinstance(dtype, torch.dtype):
    dtype = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
  dev = _native.device()
  spots_np = np.asarray(utils.to_numpy(spots), np.float64).reshape(-1)
  n_opt = spots_np.shape[0]
  expiry = _vec(expiry_times, n_opt, dtype)
  strike = _vec(strikes, n_opt, dtype) / spots_np
  vol = _vec(volatility, n_opt, dtype)
  df = _vec(discount_factors, n_opt, dtype)
  call = np.broadcast_to(np.asarray(utils.to_numpy(is_call_option),
                                    bool).reshape(-1), [n_opt])
  rate = -np.log(df) / expiry
  times = np.linspace(0.0, 1.0, int(num_exercise_times))
  # W(t) on the exercise grid, antithetic, in date-major memory [T, N]
  # (the regression kernels read one date of every path at a time): the
  # fused kernel's constant-coefficient model dX = 0 dt + 1 dW, which draws
  # the same normals per path and step as the unit GBM paths
  # exp(-t/2 + W(t)) and writes W without their exp and log
  all_times = np.concatenate([[0.0], np.asarray(times, np.float64)])
  keep = np.concatenate([[False], np.ones(len(times), bool)])
  rng = engine.RngSpec.from_random_type(RandomType.STATELESS_ANTITHETIC,
                                        seed, 0)
  w_t = engine.run(('const', 0.0, 1.0), all_times, keep, len(times),
                   np.zeros(1), int(num_samples), rng, torch.float64,
                   mode='store', dev=dev)[:, 0, :]
  w = w_t.t()  # [N, T]
  t = torch.as_tensor(times, dtype=torch.float64, device=dev).reshape(1, -1)
  t_t = t.t()
  ex = list(range(int(num_exercise_times)))
  if basis_fn is None:
    basis_fn = lsm_lib.make_polynomial_basis(2)
  bspec = getattr(basis_fn, '_tqf_basis', None)
  if bspec is not None and bspec[0] == 'poly':
    main = torch.cuda.current_stream(dev)
    streams = _streams(dev, n_opt)
    outs = []
    for b in range(n_opt):
      st = streams[b % len(streams)]
      st.wait_stream(main)  # the unit paths are written on the caller's
      with torch.cuda.stream(st):
        dfs = np.exp(-rate[b] * times * expiry[b])
        kind = (lsm_lib.PAYOFF_MAX_CALL if call[b]
                else lsm_lib.PAYOFF_BASKET_PUT)
        # ln(S / S0) in one pass: the LSM kernels take log prices
        # (basis on ln S, payoff on exp)
        log_s_t = torch.add((rate[b] - 0.5 * vol[b]**2) * (t_t * expiry[b]),
                            w_t, alpha=float(vol[b] * np.sqrt(expiry[b])))
        # [N, T, 1] view of [T][1][N] memory
        paths = log_s_t.unsqueeze(1).permute(2, 0, 1)
        outs.append(lsm_lib._native_lsm(
            paths, ex, (kind, np.array([strike[b]])), bspec[1],
            np.concatenate([[1.0], dfs]).reshape(-1, 1, 1),
            num_calibration_samples, torch.float64, dev, basis_log=2))
    for st in streams:
      main.wait_stream(st)
    prices = torch.cat(outs).cpu().numpy()
    return torch.as_tensor(prices * spots_np, dtype=dtype, device=dev)
  prices = []
  for b in range(n_opt):
    dfs = np.exp(-rate[b] * times * expiry[b])
    log_s = ((vol[b] * np.sqrt(expiry[b])) * w +
             (rate[b] - 0.5 * vol[b]**2) * (t * expiry[b]))
    k, is_call = strike[b], bool(call[b])

    def payoff_fn(log_paths, time_index, k=k, is_call=is_call):
      s = torch.exp(log_paths[:, time_index, 0])
      v = torch.relu(s - k) if is_call else torch.relu(k - s)
      return v.unsqueeze(-1)
    p = lsm_lib.least_square_mc(log_s.unsqueeze(-1), ex, payoff_fn,
                                basis_fn, discount_factors=dfs,
                                num_calibration_samples=num_calibration_samples,
                                dtype=torch.float64)
    prices.append(float(p.reshape(-1)[0]))
  out = torch.as_tensor(np.asarray(prices) * spots_np, dtype=dtype, device=dev)
  return out
