"""Euler-Maruyama sampling (mirror of tff.models.euler_sampling.sample).

Same signature and semantics as tff/models/euler_sampling.py:27-332.
Coefficient callables produced by the tqf model classes (GBM, Heston,
local volatility) run in one fused HIP kernel (tqf_euler) with the state
in registers.  When the initial state or a parameter of a GBM, constant-
coefficient or Heston model needs gradients (`watch_params`, or tensors
that require grad), the fused kernel samples and the forward-tangent
kernel (tqf_euler_tangent) differentiates, with memory flat in the number
of steps (tqf.engine.differentiable).  Any other callable pair runs as a
per-step loop of torch ops on the same GPU, fed by the engine's HIP normal
generators; autograd through that loop serves `watch_params` there.
"""
import numpy as np
import torch

from tqf import _native
from tqf import errors
from tqf.engine import differentiable
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
  batch = ()
  if normal_draws is not None:
    draws = utils.device_tensor(normal_draws, dtype, dev)
    if draws.dim() < 3:
      raise ValueError('`normal_draws` must have shape batch_shape + '
                       '[num_samples, num_time_steps, dim]')
    if draws.shape[-1] != dim:
      raise ValueError(
          '`dim` should be equal to `normal_draws.shape[2]` but are '
          '{0} and {1} respectively'.format(dim, draws.shape[-1]))
    if validate_args and draws.shape[-2] != steps:
      raise errors.InvalidArgumentError(
          '`num_time_steps` should be equal to `tf.shape(normal_draws)[1]`')
    batch = tuple(draws.shape[:-3])
    num_samples = draws.shape[-3]
  num_samples = int(num_samples)

  if initial_state is None:
    initial_state = np.zeros(dim, npd)
  if torch.is_tensor(initial_state):
    x0 = utils.device_tensor(initial_state, dtype, dev)
  else:
    # host values stay on the host: the fused launch uploads them with its
    # other small arrays in one asynchronous copy
    x0 = np.asarray(utils.to_numpy(initial_state), npd)

  def on_dev(a):
    return a if torch.is_tensor(a) else torch.as_tensor(a, device=dev)
  # batch_shape = shape(initial_state)[:-2] (euler_sampling.py:251), or the
  # leading axes of batched normal_draws
  if x0.ndim > 2:
    batch = tuple(torch.broadcast_shapes(tuple(x0.shape[:-2]), batch))

  spec = engine.spec_of(drift_fn, volatility_fn)
  fused = spec is not None and engine.model_dim(spec) == dim
  if fused and batch:
    # a batch of initial states (or of draws) for one model: the batch
    # members are consecutive blocks of paths of a single launch
    nb = int(np.prod(batch))
    xb = (on_dev(x0) + torch.zeros(batch + (num_samples, dim), dtype=dtype,
                                   device=dev)).reshape(nb * num_samples, dim)
    if draws is not None:
      draws = draws.expand(batch + tuple(draws.shape[-3:])).reshape(
          (nb * num_samples,) + tuple(draws.shape[-2:]))
    out = _fused(dim, spec, all_times, keep_mask, k, xb, nb * num_samples,
                 random_type, seed, skip, draws, watch_params, dtype, dev,
                 steps)
    if out is not None:
      return out.reshape(batch + (num_samples, k, dim))
  elif fused:
    out = _fused(dim, spec, all_times, keep_mask, k, x0, num_samples,
                 random_type, seed, skip, draws, watch_params, dtype, dev,
                 steps)
    if out is not None:
      return out
  return _generic_loop(dim, drift_fn, volatility_fn, all_times, keep_mask, k,
                       on_dev(x0), num_samples, random_type, seed, skip, draws,
                       dtype, dev, batch)


def _fused(dim, spec, all_times, keep_mask, k, x0, num_samples, random_type,
           seed, skip, draws, watch_params, dtype, dev, steps):
  """The fused kernels for a model spec; None when the call needs the
  per-step loop (watch_params on a model without tangents)."""
  if draws is None and random_type in (RandomType.HALTON,
                                       RandomType.HALTON_RANDOMIZED):
    # Halton draws are produced up front ([N, steps, d], the reference's
    # quasi layout) and read by the fused kernel as injected draws
    draws = utils.generate_mc_normal_draws(
        num_normal_draws=dim, num_time_steps=steps,
        num_sample_paths=num_samples, random_type=random_type, seed=seed,
        skip=skip, dtype=dtype).permute(1, 0, 2)
  grads = (differentiable.needs_gradients(spec, x0) or
           (watch_params is not None and
            spec[0] in differentiable.TANGENT_KINDS))
  if grads:
    # fused forward, forward-tangent backward (memory flat in steps);
    # Sobol draws are generated up front so both passes read the same
    if draws is None and random_type not in (
        None, RandomType.PSEUDO, RandomType.STATELESS,
        RandomType.PSEUDO_ANTITHETIC, RandomType.STATELESS_ANTITHETIC):
      draws = utils.generate_mc_normal_draws(
          num_normal_draws=dim, num_time_steps=steps,
          num_sample_paths=num_samples, random_type=random_type, seed=seed,
          skip=skip, dtype=dtype).permute(1, 0, 2)
    rng = engine.RngSpec.from_random_type(random_type, seed, skip, draws)
    if not torch.is_tensor(x0):
      x0 = torch.as_tensor(x0, device=dev)
    return differentiable.sample_paths(spec, all_times, keep_mask, k, x0,
                                       num_samples, rng, dtype, dev)
  if watch_params is not None:
    return None
  rng = engine.RngSpec.from_random_type(random_type, seed, skip, draws)
  # the [N, k, d] result is a view of the coalesced [k, d, N] path store
  return engine.run(spec, all_times, keep_mask, k, x0, num_samples, rng,
                    dtype, mode='view', dev=dev)


def _generic_loop(dim, drift_fn, volatility_fn, all_times, keep_mask, k, x0,
                  num_samples, random_type, seed, skip, draws, dtype, dev,
                  batch=()):
  """`_while_loop` of euler_sampling.py:405-464 as torch ops on the GPU;
  the state keeps its batch axes, batch + [N, d]."""
  steps = all_times.shape[0] - 1
  if draws is None:
    draws = utils.generate_mc_normal_draws(
        num_normal_draws=dim, num_time_steps=steps,
        num_sample_paths=num_samples, random_type=random_type, seed=seed,
        skip=skip, dtype=dtype, batch_shape=batch)  # [steps] + batch + [N, d]
  else:
    r = draws.dim()
    draws = draws.permute((r - 2,) + tuple(range(r - 2)) + (r - 1,))
  times_t = torch.as_tensor(all_times, dtype=dtype, device=dev)
  dt = times_t[1:] - times_t[:-1]
  sqrt_dt = torch.sqrt(dt)
  x = x0 + torch.zeros(batch + (num_samples, dim), dtype=dtype, device=dev)
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
  zero = torch.zeros_like(x)
  return torch.stack([o if o is not None else zero for o in outs], dim=-2)
