"""Host-side set-up shared by the samplers (mirror of tff/models/utils.py).

  prepare_grid / _grid_from_time_step / _grid_from_num_times
      tff/models/utils.py:209-320 (time grid, keep mask, time indices)
  generate_mc_normal_draws
      tff/models/utils.py:20-128 (draws [steps, N, d]; quasi dimension of
      step i, component k is i*d + k)
  record_plan
      turns the `_while_loop` record rule of euler_sampling.py:405-464 into a
      per-step slot table for the kernels.

Grids are a few hundred numbers built once per call on the host, as the
reference builds them at graph-construction time.
"""
import numpy as np
import torch

from tqf import _native
from tqf.math.random import multivariate_normal as mvn

__all__ = ['prepare_grid', 'record_plan', 'generate_mc_normal_draws',
           'to_numpy', 'np_dtype']


def np_dtype(dtype):
  if dtype in (torch.float64, np.float64):
    return np.float64
  if dtype in (torch.float32, np.float32):
    return np.float32
  return np.dtype(dtype).type


def scalar_param(x):
  """A scalar model parameter: a tensor that requires grad stays a tensor
  (the engine can differentiate with respect to it); anything else becomes
  a float."""
  if isinstance(x, torch.Tensor) and x.requires_grad:
    if x.numel() != 1:
      raise NotImplementedError('batched model parameters are not supported')
    return x.reshape(())
  return float(to_numpy(x))


def device_tensor(x, dtype, dev):
  """x on the device in `dtype`.  Python numbers and lists are converted in
  `dtype` directly (torch would build them in float32 first and round a
  float64 input); a tensor keeps its autograd history."""
  if isinstance(x, torch.Tensor):
    return x.to(device=dev, dtype=dtype)
  return torch.as_tensor(np.asarray(x, dtype=np_dtype(dtype)), device=dev)


def to_numpy(x, dtype=None):
  if isinstance(x, torch.Tensor):
    x = x.detach().cpu().numpy()
  a = np.asarray(x)
  return a.astype(dtype) if dtype is not None else a


def _grid_from_time_step(times, time_step, tolerance, dtype):
  grid = np.arange(0.0, float(times[-1]), float(time_step)).astype(dtype)
  all_times = np.sort(np.concatenate([times, grid]).astype(dtype),
                      kind='stable')
  dt = np.concatenate([np.ones(1, dtype), all_times[1:] - all_times[:-1]])
  all_times = all_times[dt > tolerance]
  idx = np.searchsorted(all_times, times, side='left')
  idx = np.minimum(idx, all_times.shape[0] - 1)
  idx = np.where(all_times[idx] - times > tolerance, idx - 1, idx)
  return all_times, idx


def _grid_from_num_times(times, time_step, num_time_steps, dtype):
  n_uniform = max(int(num_time_steps) - times.shape[0], 0)
  uniform = np.linspace(time_step, times[-1] - time_step, n_uniform,
                        dtype=dtype)
  grid = np.sort(np.concatenate([uniform, times]).astype(dtype),
                 kind='stable')
  all_times = np.concatenate([np.zeros(1, dtype), grid]).astype(dtype)
  return all_times, np.searchsorted(all_times, times, side='left')


def prepare_grid(*, times, time_step, dtype, tolerance=None,
                 num_time_steps=None, times_grid=None):
  """Returns (all_times, keep_mask, time_indices) as numpy arrays."""
  dtype = np_dtype(dtype)
  times = to_numpy(times, dtype)
  if tolerance is None:
    tolerance = 1e-10 if dtype == np.float64 else 1e-6
  tolerance = dtype(tolerance)
  if times_grid is None:
    if num_time_steps is None:
      all_times, idx = _grid_from_time_step(times, dtype(time_step),
                                            tolerance, dtype)
    else:
      all_times, idx = _grid_from_num_times(times, dtype(time_step),
                                            num_time_steps, dtype)
  else:
    all_times = to_numpy(times_grid, dtype)
    idx = np.searchsorted(all_times, times, side='left')
    idx = np.minimum(idx, all_times.shape[0] - 1)
    d1 = all_times[idx] - times
    d2 = all_times[np.maximum(idx - 1, 0)] - times
    idx = np.where(np.abs(d2) > np.abs(d1), idx, np.maximum(idx - 1, 0))
  mask = np.zeros(all_times.shape[0], bool)
  mask[idx] = True
  return all_times, mask, idx.astype(np.int64)


def record_plan(keep_mask, num_requested, initial_written=None):
  """Per-step write slots of the reference's record loop.

  The loop writes the state after step i at `written_count` and then adds
  keep_mask[i+1]; it stops once `written_count` reaches the number of
  requested times.  A slot keeps the last state written to it, so only the
  final writer of each slot matters.  Returns (slot[steps_executed] with -1
  for overwritten writes, init_slot, steps_executed, slots_filled).
  `initial_written` overrides the starting count (HestonModel starts its
  loop at 0 whatever keep_mask[0] is, heston_model.py:443-446).
  """
  keep = np.asarray(keep_mask, bool)
  steps = keep.shape[0] - 1
  k = int(num_requested)
  record = k != 1
  written0 = int(keep[0]) if initial_written is None else int(initial_written)
  # the count at the top of step i: written0 + keep[1] + ... + keep[i]
  cnt = written0 + np.concatenate(
      [[0], np.cumsum(keep[1:max(steps, 1)], dtype=np.int64)])[:steps]
  stop = np.flatnonzero(cnt >= k)
  executed = int(stop[0]) if stop.size else steps
  slot = cnt[:executed] if record else np.zeros(executed, np.int64)
  # the count never decreases, so a slot's writes are consecutive steps and
  # its final writer is the last of them
  last = np.ones(executed, bool)
  last[:-1] = slot[1:] != slot[:-1]
  slots = np.where(last, slot, -1).astype(np.int32)
  # slot 0 starts as the initial state unless a step writes it
  init_slot = -1 if executed and slot[0] == 0 else 0
  filled = sorted({0} | set(slot.tolist()))
  return slots, init_slot, executed, filled


def generate_mc_normal_draws(num_normal_draws, num_time_steps,
                             num_sample_paths, random_type, batch_shape=None,
                             skip=0, seed=None, dtype=None, name=None):
  """Draws [num_time_steps] + batch_shape + [N, d], laid out as the
  reference lays them out (models/utils.py:20-128); batch member b takes the
  stream's paths [b N, (b + 1) N)."""
  del name
  batch = tuple(int(b) for b in (batch_shape or ()))
  if batch:
    nb = int(np.prod(batch))
    flat = generate_mc_normal_draws(num_normal_draws, num_time_steps,
                                    nb * int(num_sample_paths), random_type,
                                    skip=skip, seed=seed, dtype=dtype)
    steps, d = flat.shape[0], flat.shape[-1]
    return flat.reshape((steps,) + batch + (int(num_sample_paths), d))
  dtype = dtype or torch.float32
  if isinstance(dtype, np.dtype) or dtype in (np.float32, np.float64):
    dtype = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
  d = int(num_normal_draws)
  steps = int(num_time_steps)
  n = int(num_sample_paths)
  if random_type is None:
    random_type = mvn.RandomType.PSEUDO
  if random_type == mvn.RandomType.SOBOL:
    flat = mvn.sobol_normals(n, steps * d, int(skip or 0), dtype)
  elif random_type in (mvn.RandomType.PSEUDO, mvn.RandomType.STATELESS,
                       mvn.RandomType.PSEUDO_ANTITHETIC,
                       mvn.RandomType.STATELESS_ANTITHETIC):
    anti = random_type in (mvn.RandomType.PSEUDO_ANTITHETIC,
                           mvn.RandomType.STATELESS_ANTITHETIC)
    if anti and n % 2:
      raise ValueError('First dimension of `sample_shape` should be even for '
                       'PSEUDO_ANTITHETIC random type')
    flat = mvn.philox_normals(n, steps * d, mvn.seed_to_key(seed, random_type),
                              dtype, antithetic=anti)
  elif random_type in (mvn.RandomType.HALTON,
                       mvn.RandomType.HALTON_RANDOMIZED):
    # quasi draws: sequence index skip + p, dimension step * d + k
    from tqf.math.random import halton
    flat, _ = halton.sample(
        steps * d, sequence_indices=np.arange(int(skip or 0),
                                              int(skip or 0) + n),
        randomized=random_type == mvn.RandomType.HALTON_RANDOMIZED,
        seed=seed, dtype=dtype, _normal=True)
  else:
    raise NotImplementedError('random type {} is not supported'.format(
        random_type))
  return flat.reshape(n, steps, d).permute(1, 0, 2)


def heston_grid(times, time_step, num_time_steps, extra_times, dtype):
  """HestonModel grid (heston_model.py:575-639): grid U times U jumps,
  stable argsort, duplicates kept.  Returns (all_times, keep_mask)."""
  dtype = np_dtype(dtype)
  times = to_numpy(times, dtype).reshape(-1)
  if num_time_steps is not None:
    grid = np.linspace(dtype(0.0), times[-1] - dtype(time_step),
                       int(num_time_steps), dtype=dtype)
  else:
    grid = np.arange(0.0, float(times[-1]), float(time_step)).astype(dtype)
  parts = [grid, times] + [np.asarray(e, dtype) for e in extra_times]
  all_times = np.concatenate(parts).astype(dtype)
  mask = np.concatenate([np.zeros(grid.shape[0], bool),
                         np.ones(times.shape[0], bool)] +
                        [np.zeros(np.asarray(e).shape[0], bool)
                         for e in extra_times])
  perm = np.argsort(all_times, kind='stable')
  return all_times[perm], mask[perm]
