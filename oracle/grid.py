"""Oracle: simulation time grids.

  * `prepare_grid`           tff/models/utils.py:209-280
  * `_grid_from_time_step`   tff/models/utils.py:283-306
  * `_grid_from_num_times`   tff/models/utils.py:309-320
  * `heston_prepare_grid`    tff/models/heston/heston_model.py:575-639

Pinned by: grid sizes and contents that follow from the definitions
(num_time_steps=100, times=[1] gives [0, .01, ..., .99, 1] up to rounding;
365 steps with 4 maturities gives 366 points containing every maturity);
record masks cross-checked by searchsorted on the result; duplicates removed
within tolerance (euler_sampling_test.py grid cases).
"""
import numpy as np


def _grid_from_num_times(times, time_step, num_time_steps, dtype):
  n_uniform = max(num_time_steps - times.shape[0], 0)
  uniform = np.linspace(dtype(time_step), dtype(times[-1] - time_step),
                        n_uniform).astype(dtype)
  grid = np.sort(np.concatenate([uniform, times]).astype(dtype), kind='stable')
  all_times = np.concatenate([np.zeros(1, dtype), grid]).astype(dtype)
  idx = np.searchsorted(all_times, times, side='left')
  return all_times, idx


def _grid_from_time_step(times, time_step, tolerance, dtype):
  grid = np.arange(0.0, float(times[-1]), float(time_step)).astype(dtype)
  all_times = np.sort(np.concatenate([times, grid]).astype(dtype),
                      kind='stable')
  dt = np.concatenate([[dtype(1.0)], all_times[1:] - all_times[:-1]])
  all_times = all_times[dt > tolerance]
  idx = np.searchsorted(all_times, times, side='left')
  idx = np.minimum(idx, all_times.shape[0] - 1)
  idx = np.where(all_times[idx] - times > tolerance, idx - 1, idx)
  return all_times, idx


def prepare_grid(times, dtype, time_step=None, num_time_steps=None,
                 times_grid=None, tolerance=None):
  """Returns (all_times, keep_mask, time_indices)."""
  times = np.asarray(times, dtype=dtype)
  if tolerance is None:
    tolerance = 1e-10 if dtype == np.float64 else 1e-6
  if times_grid is None:
    if num_time_steps is None:
      all_times, idx = _grid_from_time_step(times, dtype(time_step),
                                            dtype(tolerance), dtype)
    else:
      if time_step is None:
        time_step = times[-1] / dtype(num_time_steps)
      all_times, idx = _grid_from_num_times(times, dtype(time_step),
                                            int(num_time_steps), dtype)
  else:
    all_times = np.asarray(times_grid, dtype=dtype)
    idx = np.searchsorted(all_times, times, side='left')
    idx_c = np.minimum(idx, all_times.shape[0] - 1)
    d1 = all_times[idx_c] - times
    d2 = all_times[np.maximum(idx_c - 1, 0)] - times
    idx = np.where(np.abs(d2) > np.abs(d1), idx_c, np.maximum(idx_c - 1, 0))
  mask = np.zeros(all_times.shape[0], dtype=bool)
  mask[idx] = True
  return all_times, mask, idx.astype(np.int64)


def heston_prepare_grid(times, dtype, time_step=None, num_time_steps=None):
  """HestonModel grid: linspace(0, T - dt, n) U times, stable argsort.

  Duplicates are kept (heston_model.py:613-635).
  """
  times = np.asarray(times, dtype=dtype)
  if time_step is None:
    time_step = times[-1] / dtype(num_time_steps)
  if num_time_steps is None:
    grid = np.arange(0.0, float(times[-1]), float(time_step)).astype(dtype)
  else:
    grid = np.linspace(dtype(0.0), dtype(times[-1] - time_step),
                       int(num_time_steps)).astype(dtype)
  all_times = np.concatenate([grid, times]).astype(dtype)
  mask = np.concatenate([np.zeros(grid.shape[0], bool),
                         np.ones(times.shape[0], bool)])
  perm = np.argsort(all_times, kind='stable')
  return all_times[perm], mask[perm]
