"""Fused European pricing: simulate and reduce payoffs in one kernel.

The reference prices by sampling paths with `euler_sampling.sample` and
averaging a payoff in user code (Monte_Carlo_Euler_Scheme.ipynb:273-274).
Here the payoff reduction happens inside the stepper (TQF_OUT_PAYOFF), so
no path ever leaves registers.
"""
import numpy as np
import torch

from tqf import _native
from tqf.engine import euler as engine
from tqf.math import qmc
from tqf.models import utils

__all__ = ['european_calls', 'heston_ft_calls', 'sobol_digital_shift',
           'heston_grid_for']

SHIFT_STREAM = 0x5348


def sobol_digital_shift(ndims, seed):
  """32-bit digital shift words of the engine (Philox, sub-stream 0x5348)."""
  return qmc.philox_words(int(ndims), int(seed), stream_id=SHIFT_STREAM)


def heston_grid_for(maturities, num_time_steps, grid, dtype):
  """('reference': linspace grid of models/utils.py:309-320 through the
  maturities; 'uniform': T/num_time_steps spacing, maturities snapped)."""
  npd = utils.np_dtype(dtype)
  mats = np.asarray(maturities, np.float64)
  if grid == 'reference':
    all_times, keep, _ = utils.prepare_grid(
        times=mats.astype(npd), time_step=npd(mats[-1] / num_time_steps),
        num_time_steps=num_time_steps, dtype=npd)
    return all_times, keep
  if grid == 'uniform':
    dt = mats[-1] / num_time_steps
    all_times = (np.arange(num_time_steps + 1) * dt).astype(npd)
    idx = np.rint(mats / dt).astype(np.int64)
    keep = np.zeros(num_time_steps + 1, bool)
    keep[idx] = True
    return all_times, keep
  raise ValueError('grid must be "reference" or "uniform"')


def european_calls(spec, x0, strikes, all_times, keep_mask, num_paths,
                   rng, discount_rate=0.0, dtype=torch.float32, prepare=False):
  """Discounted call prices and standard errors at the recorded times.

  spec: engine coefficient spec; log-models price on exp(x[0]).
  Returns dict(prices [k, K], stderrs [k, K]) as float64 tensors, or the
  prepared engine.Launch when `prepare` is true.
  """
  keep = np.asarray(keep_mask, bool)
  times = np.asarray(all_times, np.float64)[keep]
  disc = np.exp(-float(discount_rate) * times)
  launch = engine.prepare(spec, all_times, keep, int(keep.sum()), x0,
                          int(num_paths), rng, dtype, mode='payoff',
                          strikes=strikes, discount=disc)
  return launch if prepare else launch.launch()


def heston_ft_calls(s0=100.0, v0=0.04, kappa=2.0, theta=0.04, xi=0.3,
                    rho=-0.7, r=0.03, strikes=None, maturities=(0.25, 0.5,
                                                                0.75, 1.0),
                    num_paths=2**20, num_time_steps=365, grid='reference',
                    random='sobol_shifted', seed=7, dtype=torch.float32,
                    prepare=False):
  """Full-truncation log-Euler Heston calls over strikes x maturities."""
  if strikes is None:
    strikes = np.linspace(60.0, 140.0, 64)
  all_times, keep = heston_grid_for(maturities, num_time_steps, grid, dtype)
  steps = all_times.shape[0] - 1
  if random == 'sobol_shifted':
    rng = engine.RngSpec('sobol_shifted',
                         shift=sobol_digital_shift(2 * steps, seed),
                         first_index=0)
  elif random == 'philox':
    rng = engine.RngSpec('philox', seed=seed)
  else:
    raise ValueError('random must be "sobol_shifted" or "philox"')
  spec = ('heston', r, kappa, theta, xi, rho, 0.0)
  x0 = np.array([np.log(s0), v0])
  return european_calls(spec, x0, strikes, all_times, keep, num_paths, rng,
                        discount_rate=r, dtype=dtype, prepare=prepare)
