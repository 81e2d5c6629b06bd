"""Euler-Maruyama sampling (mirror of tff.models.euler_sampling.sample).

Same signature and semantics as tff/models/euler_sampling.py:27-332.
Coefficient callables produced by the tqf model classes (GBM, Heston,
local volatility) run in one fused HIP kernel (tqf_euler) with the state
in registers.  Any other callable pair runs as a per-step loop of torch
ops on the same GPU, fed by the engine's HIP normal generators; that loop is
also the differentiable route for `watch_params`.
"""
import numpy as np
import torch

from tqf import _native
from tqf import errors
from tqf.engine import euler as engine
from tqf.math.random.multivariate_normal import RandomType
from tqf.models import utils

__all__ = ['sample']


def _resolve_dtype(dtype, times):
  if dtype is not None:
    if isinstance(dtype, torch.dtype):
      return dtype
    return torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
  if isinstance(times, torch.Tensor):
    return times.dtype
  if isinstance(times, np.ndarray) and times.dtype == np.float64:
    return torch.float64
  return torch.float32


def sample(dim, drift_fn, volatility_fn, times, time_step=None,
           num_time_steps=None, num_samples=1, initial_state=None,
           random_type=None, seed=None, swap_memory=True, skip=0,
           precompute_normal_draws=True, times_grid=None, normal_draws=None,
           watch_params=None, validate_args=False, tolerance=None,
           dtype=None, name=None):
  """Returns sample paths [num_samples, k, dim] of dX = a dt + b dW."""
  del swap_memory, precompute_normal_draws, name
  dtype = _resolve_dtype(dtype, times)
  npd = utils.np_dtype(dtype)
  times_np = utils.to_numpy(times, npd).reshape(-1)
  if tolerance is None:
    tolerance = 1e-10 if dtype == torch.float64 else 1e-6
  if validate_args and np.any(times_np[1:] <= times_np[:-1] + tolerance):
    raise errors.InvalidArgumentError(
        '`times` increments should be greater than tolerance {0}'.format(
            tolerance))
  if num_time_steps is not None and time_step is not None:
    raise ValueError(
        'When `times_grid` is not supplied only one of either '
        '`num_time_steps` or `time_step` should be defined but not both.')
  if times_grid is None:
    if time_step is None:
      if num_time_steps is None:
        raise ValueError(
            'When `times_grid` is not supplied, either `num_time_steps` '
            'or `time_step` should be defined.')
      time_step = times_np[-1] / npd(num_time_steps)
  elif validate_args:
    tg = utils.to_numpy(times_grid, npd)
    if np.any(tg[1:] <= tg[:-1] + tolerance):
      raise errors.InvalidArgumentError(
          '`times_grid` increments should be greater than tolerance '
          '{0}'.format(tolerance))
  all_times, keep_mask, _ = utils.prepare_grid(
      times=times_np, time_step=time_step, num_time_steps=num_time_steps,
      times_grid=times_grid, tolerance=tolerance, dtype=npd)
  steps = all_times.shape[0] - 1
  k = times_np.shape[0]

  dev = _native.device()
  draws = None
  if normal_draws is not None:
    draws = torch.as_tensor(normal_draws, device=dev).to(dtype)
    if draws.dim() != 3:
      raise NotImplementedError('batched normal_draws are not supported')
    if draws.shape[-1] != dim:
      raise ValueError(
          '`dim` should be equal to `normal_draws.shape[2]` but are '
          '{0} and {1} respectively'.format(dim, draws.shape[-1]))
    if validate_args and draws.shape[1] != steps:
      raise errors.InvalidArgumentError(
          '`num_time_steps` should be equal to `tf.shape(normal_draws)[1]`')
    num_samples = draws.shape[0]
  num_samples = int(num_samples)

  if initial_state is None:
    initial_state = torch.zeros(dim, dtype=dtype, device=dev)
  x0 = torch.as_tensor(initial_state, device=dev).to(dtype)
  if x0.dim() > 2:
    raise NotImplementedError('batch_shape in initial_state is not supported')

  spec = engine.spec_of(drift_fn, volatility_fn)
  if (spec is not None and draws is None and random_type in
      (RandomType.HALTON, RandomType.HALTON_RANDOMIZED)):
    # Halton draws are produced up front ([N, steps, d], the reference's
    # quasi layout) and read by the fused kernel as injected draws
    draws = utils.generate_mc_normal_draws(
        num_normal_draws=dim, num_time_steps=steps,
        num_sample_paths=num_samples, random_type=random_type, seed=seed,
        skip=skip, dtype=dtype).permute(1, 0, 2)
  if spec is not None and watch_params is None and engine.model_dim(spec) == dim:
    rng = engine.RngSpec.from_random_type(random_type, seed, skip, draws)
    return engine.run(spec, all_times, keep_mask, k, x0, num_samples, rng,
                      dtype, mode='paths', dev=dev)
  return _generic_loop(dim, drift_fn, volatility_fn, all_times, keep_mask, k,
                       x0, num_samples, random_type, seed, skip, draws, dtype,
                       dev)


def _generic_loop(dim, drift_fn, volatility_fn, all_times, keep_mask, k, x0,
                  num_samples, random_type, seed, skip, draws, dtype, dev):
  """`_while_loop` of euler_sampling.py:405-464 as torch ops on the GPU."""
  steps = all_times.shape[0] - 1
  if draws is None:
    draws = utils.generate_mc_normal_draws(
        num_normal_draws=dim, num_time_steps=steps,
        num_sample_paths=num_samples, random_type=random_type, seed=seed,
        skip=skip, dtype=dtype)  # [steps, N, d]
  else:
    draws = draws.permute(1, 0, 2)
  times_t = torch.as_tensor(all_times, dtype=dtype, device=dev)
  dt = times_t[1:] - times_t[:-1]
  sqrt_dt = torch.sqrt(dt)
  x = x0 + torch.zeros((num_samples, dim), dtype=dtype, device=dev)
  slots, init_slot, executed, _ = utils.record_plan(keep_mask, k)
  outs = [None] * k
  if init_slot >= 0:
    outs[init_slot] = x
  for i in range(executed):
    t = times_t[i + 1]
    dw = draws[i] * sqrt_dt[i]
    inc = dt[i] * drift_fn(t, x)
    vol = volatility_fn(t, x)
    x = x + inc + torch.matmul(vol, dw.unsqueeze(-1)).squeeze(-1)
    if slots[i] >= 0:
      outs[slots[i]] = x
  zero = torch.zeros((num_samples, dim), dtype=dtype, device=dev)
  return torch.stack([o if o is not None else zero for o in outs], dim=1)
