"""Normal draws by random type (mirror of tff.math.random.mv_normal_sample).

Follows tff/math/random_ops/multivariate_normal.py:27-451.  The pseudo
random types use the engine's Philox4x32-10 + Box-Muller stream (the TF
Philox stream itself cannot be reproduced without TF); SOBOL keeps the
reference's layout and transform bit-for-bit up to the erfinv evaluation:
points of sobol.sample -> erfinv((u - .5) * 2) * sqrt(2) -> the
reverse-batch reshape of `_mvnormal_quasi` (multivariate_normal.py:402-420).
All draws are produced by HIP kernels.
"""
import ctypes
import enum
import random as _pyrandom

import numpy as np
import torch

from tqf import _native
from tqf.math.random import sobol as sobol_lib

__all__ = ['RandomType', 'multivariate_normal', 'philox_normals',
           'sobol_normals', 'seed_to_key']


@enum.unique
class RandomType(enum.Enum):
  """Types of random number sequences (multivariate_normal.py:27-44)."""
  PSEUDO = 0
  STATELESS = 1
  HALTON = 2
  HALTON_RANDOMIZED = 3
  SOBOL = 4
  PSEUDO_ANTITHETIC = 5
  STATELESS_ANTITHETIC = 6


def seed_to_key(seed, random_type=None):
  """Philox key (64-bit) from an int seed or a [2] stateless seed."""
  if random_type in (RandomType.STATELESS, RandomType.STATELESS_ANTITHETIC):
    if seed is None:
      raise ValueError('`seed` should be specified if the `random_type` is '
                       '`STATELESS` or `STATELESS_ANTITHETIC`')
  if seed is None:
    return _pyrandom.getrandbits(64)
  if isinstance(seed, torch.Tensor):
    seed = seed.detach().cpu().numpy()
  if np.ndim(seed) == 1:
    s = np.asarray(seed).astype(np.int64)
    return (int(s[0]) & 0xFFFFFFFF) | ((int(s[1]) & 0xFFFFFFFF) << 32)
  return int(seed) & 0xFFFFFFFFFFFFFFFF


def philox_normals(n_paths, n_draws, key, dtype, antithetic=False,
                   stream_id=0, dev=None):
  """[n_paths, n_draws] engine-stream normals (tqf_philox_normals)."""
  dev = dev or _native.device()
  out = torch.empty((n_paths, n_draws), dtype=dtype, device=dev)
  if n_paths and n_draws:
    _native.call('tqf_philox_normals', ctypes.c_uint64(key),
                 ctypes.c_uint32(stream_id), ctypes.c_int64(n_paths),
                 ctypes.c_int64(n_draws), ctypes.c_int32(int(antithetic)),
                 ctypes.c_int32(_native.dtype_code(dtype)), _native.ptr(out),
                 _native.stream_ptr(dev))
  return out


def sobol_normals(num_samples, dim, skip, dtype, dev=None):
  """[N, dim] normals of sobol.sample(dim, N, skip) (reference transform)."""
  dev = dev or _native.device()
  nd = sobol_lib.num_digits_for(skip, num_samples)
  v = sobol_lib.device_columns(dim, nd, dev=dev)
  out = torch.empty((num_samples, dim), dtype=dtype, device=dev)
  if num_samples:
    _native.call('tqf_sobol_normals', _native.ptr(v), ctypes.c_int32(dim),
                 ctypes.c_int32(nd), ctypes.c_int64(num_samples),
                 ctypes.c_int64(skip + 1), None, ctypes.c_int32(0),
                 ctypes.c_int32(_native.dtype_code(dtype)), _native.ptr(out),
                 _native.stream_ptr(dev))
  return out


def _as_tensor(x, dtype, dev):
  if x is None:
    return None
  if isinstance(x, torch.Tensor):
    return x.to(device=dev, dtype=dtype)
  return torch.as_tensor(np.asarray(x), dtype=dtype, device=dev)


def _resolve_dtype(dtype, *tensors):
  if dtype is not None:
    return sobol_lib._torch_dtype(dtype)
  for t in tensors:
    if isinstance(t, torch.Tensor) and t.is_floating_point():
      return t.dtype
    if isinstance(t, np.ndarray) and t.dtype.kind == 'f':
      return sobol_lib._torch_dtype(t.dtype)
  return torch.float32


def multivariate_normal(sample_shape, mean=None, covariance_matrix=None,
                        scale_matrix=None, random_type=None,
                        validate_args=False, seed=None, dtype=None, name=None,
                        **kwargs):
  """Draws from N(mean, covariance) by `random_type`.

  Output shape sample_shape + batch_shape where batch_shape is mean's shape
  (or the scale matrix's shape without its last axis).
  """
  del validate_args, name
  if covariance_matrix is not None and scale_matrix is not None:
    raise ValueError('Only one of covariance matrix or scale matrix'
                     ' must be specified')
  if random_type is None:
    random_type = RandomType.PSEUDO
  dev = _native.device()
  dtype = _resolve_dtype(dtype, mean, scale_matrix, covariance_matrix)
  sample_shape = [int(s) for s in np.atleast_1d(np.asarray(sample_shape))]
  mean_t = _as_tensor(mean, dtype, dev)
  scale = _as_tensor(scale_matrix, dtype, dev)
  if scale is None and covariance_matrix is not None:
    cov = _as_tensor(covariance_matrix, torch.float64, dev)
    scale = torch.linalg.cholesky(cov).to(dtype)
  if mean_t is None:
    if scale is None:
      raise ValueError('At least one of mean, covariance_matrix or '
                       'scale_matrix must be specified')
    batch_shape = list(scale.shape[:-1])
  else:
    batch_shape = list(mean_t.shape)
  dim = batch_shape[-1]
  n_rows = int(np.prod(sample_shape)) if sample_shape else 1
  n_cols = int(np.prod(batch_shape))

  if random_type in (RandomType.PSEUDO, RandomType.STATELESS,
                     RandomType.PSEUDO_ANTITHETIC,
                     RandomType.STATELESS_ANTITHETIC):
    anti = random_type in (RandomType.PSEUDO_ANTITHETIC,
                           RandomType.STATELESS_ANTITHETIC)
    if anti and sample_shape[0] % 2:
      raise ValueError('First dimension of `sample_shape` should be even for '
                       'PSEUDO_ANTITHETIC random type')
    key = seed_to_key(seed, random_type)
    z = philox_normals(n_rows, n_cols, key, dtype, antithetic=anti, dev=dev)
    z = z.reshape(sample_shape + batch_shape)
  elif random_type == RandomType.SOBOL:
    skip = int(kwargs.get('skip', 0) or 0)
    batch_rev = batch_shape[::-1]
    out_shape_t = batch_rev + sample_shape
    num_samples = int(np.prod(out_shape_t)) // dim
    zs = sobol_normals(num_samples, dim, skip, dtype, dev=dev)  # [n, dim]
    zt = zs.transpose(0, 1).reshape(out_shape_t)
    nb, ns = len(batch_shape), len(sample_shape)
    perm = list(range(nb, nb + ns)) + list(range(nb - 1, -1, -1))
    z = zt.permute(perm)
  elif random_type in (RandomType.HALTON, RandomType.HALTON_RANDOMIZED):
    # _mvnormal_quasi (multivariate_normal.py:393-424): Halton points of
    # sequence indices skip .. skip + n - 1, the reference transform
    from tqf.math.random import halton
    skip = int(kwargs.get('skip', 0) or 0)
    batch_rev = batch_shape[::-1]
    out_shape_t = batch_rev + sample_shape
    num_samples = int(np.prod(out_shape_t)) // dim
    zs, _ = halton.sample(
        dim, sequence_indices=np.arange(skip, skip + num_samples),
        randomized=random_type == RandomType.HALTON_RANDOMIZED,
        randomization_params=kwargs.get('randomization_params'), seed=seed,
        dtype=dtype, _normal=True)
    zt = zs.transpose(0, 1).reshape(out_shape_t)
    nb, ns = len(batch_shape), len(sample_shape)
    perm = list(range(nb, nb + ns)) + list(range(nb - 1, -1, -1))
    z = zt.permute(perm)
  else:
    raise NotImplementedError(
        'random type {} is not supported'.format(random_type))
  if scale is not None:
    z = torch.einsum('...ij,...j->...i', scale, z)
  if mean_t is not None:
    z = mean_t + z
  return z
