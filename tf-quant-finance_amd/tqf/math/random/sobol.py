"""Sobol sequences on the GPU (mirror of tff.math.random.sobol).

`sample` keeps the reference's semantics (tff/math/random_ops/sobol/
sobol_impl.py:39-167): natural (not Gray-code) order, point i is sequence
index skip + 1 + i, value = cast(int) / 2^nd with nd = ceil(log2(skip + N +
1)).  The integer points are bit-identical to the reference's; the XOR walk
runs in the `tqf_sobol_points` HIP kernel, one thread per (point, dim).
"""
import ctypes
import math
import os

import numpy as np
import torch

from tqf import _native
from tqf import errors

__all__ = ['sample', 'load_data', 'direction_numbers', 'generating_columns']

_DATA_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__)))), 'data', 'sobol', 'new-joe-kuo-6.21201')
_MAX_POSITIVE = 2**31 - 1

_data = None
_m_cache = {}
_dev_cache = {}


def load_data():
  """Parses the vendored Joe-Kuo file (sobol_impl.py:237-261).

  Returns (poly int64 [21200], minit int64 [21200, 18]) where
  poly = 2^s + 2a + 1 and minit holds the initial direction numbers.
  """
  global _data
  if _data is None:
    poly = np.zeros(21200, np.int64)
    minit = np.zeros((21200, 18), np.int64)
    with open(_DATA_PATH) as f:
      next(f)
      k = 0
      for line in f:
        tok = line.split()
        if not tok:
          continue
        s, a = int(tok[1]), int(tok[2])
        poly[k] = 2**s + 2 * a + 1
        vals = [int(t) for t in tok[3:3 + s]]
        minit[k, :len(vals)] = vals
        k += 1
    _data = (poly, minit)
  return _data


def direction_numbers(dim):
  """m[dim, 32] uint32 direction numbers (native host routine)."""
  if dim not in _m_cache:
    poly, minit = load_data()
    out = np.zeros((dim, 32), np.uint32)
    _native.call('tqf_sobol_direction_numbers',
                 poly.ctypes.data_as(ctypes.c_void_p),
                 minit.ctypes.data_as(ctypes.c_void_p),
                 ctypes.c_int32(poly.shape[0]), ctypes.c_int32(dim),
                 out.ctypes.data_as(ctypes.c_void_p))
    _m_cache[dim] = out
  return _m_cache[dim]


def generating_columns(dim, num_digits, num_columns=None):
  """v[d][j] = m_j << (num_digits - 1 - j) for j < num_columns, else 0."""
  if num_columns is None:
    num_columns = num_digits
  num_columns = min(num_columns, num_digits, 32)
  m = direction_numbers(dim).astype(np.uint64)
  v = np.zeros((dim, 32), np.uint64)
  for j in range(num_columns):
    v[:, j] = (m[:, j] << np.uint64(num_digits - 1 - j)) & np.uint64(
        0xFFFFFFFF)
  return v.astype(np.uint32)


def device_columns(dim, num_digits, num_columns=None, dev=None):
  """Cached device copy of `generating_columns` (int32 view of uint32)."""
  dev = dev or _native.device()
  key = (dim, num_digits, num_columns, str(dev))
  t = _dev_cache.get(key)
  if t is None:
    v = generating_columns(dim, num_digits, num_columns)
    t = torch.from_numpy(v.view(np.int32)).to(dev)
    _dev_cache[key] = t
  return t


def device_lo_table(dim, num_digits, group=1, dev=None):
  """Low-byte XOR table [dim/group][256][group] for the fused steppers."""
  dev = dev or _native.device()
  key = ('lo', dim, num_digits, group, str(dev))
  t = _dev_cache.get(key)
  if t is None:
    v = device_columns(dim, num_digits, dev=dev)
    t = torch.empty((dim // group, 256, group), dtype=torch.int32, device=dev)
    _native.call('tqf_sobol_lo_table', _native.ptr(v), ctypes.c_int32(dim),
                 ctypes.c_int32(group), _native.ptr(t),
                 _native.stream_ptr(dev))
    _dev_cache[key] = t
  return t


def num_digits_for(skip, num_results):
  """ceil(log2(skip + N + 1)) in float64 (sobol_impl.py:118-123)."""
  return int(math.ceil(math.log(float(skip + num_results + 1)) / math.log(2.)))


def _torch_dtype(dtype, default=torch.float32):
  if dtype is None:
    return default
  if isinstance(dtype, torch.dtype):
    return dtype
  return {np.dtype(np.float16): torch.float16,
          np.dtype(np.float32): torch.float32,
          np.dtype(np.float64): torch.float64}[np.dtype(dtype)]


def sample(dim, num_results, skip=0, validate_args=False, dtype=None,
           name=None):
  """Returns num_results Sobol points of dimension dim, shape [N, dim].

  Same arguments and meaning as tff.math.random.sobol.sample.  Output is a
  torch tensor on the current ROCm device.
  """
  del name
  num_results = int(num_results)
  skip = int(skip)
  if validate_args:
    if dim < 1:
      raise ValueError(
          'Dimension must be greater than zero. Supplied {}'.format(dim))
    if num_results <= 0:
      raise errors.InvalidArgumentError(
          'Number of results `num_results` must be greater than zero.')
    if skip < 0:
      raise errors.InvalidArgumentError('`skip` must be non-negative.')
    if not _MAX_POSITIVE - num_results > skip:
      raise errors.InvalidArgumentError(
          'Skip too large. Should be smaller than {} - num_results'.format(
              _MAX_POSITIVE))
    num_results = max(num_results, 1)
    skip = max(skip, 0)
  if dim < 1:
    raise ValueError(
        'Dimension must be greater than zero. Supplied {}'.format(dim))
  tdt = _torch_dtype(dtype)
  dev = _native.device()
  nd = num_digits_for(skip, num_results)
  if nd > 32:
    raise errors.InvalidArgumentError('Sobol index range exceeds 32 bits')
  v = device_columns(dim, nd, dev=dev)
  kind = {torch.float32: 1, torch.float64: 2, torch.float16: 2}[tdt]
  out_dt = torch.float64 if tdt == torch.float16 else tdt
  out = torch.empty((num_results, dim), dtype=out_dt, device=dev)
  if num_results:
    _native.call('tqf_sobol_points', _native.ptr(v), ctypes.c_int32(dim),
                 ctypes.c_int32(nd), ctypes.c_int64(num_results),
                 ctypes.c_int64(skip + 1), None, None, ctypes.c_int32(kind),
                 ctypes.c_int32(0), _native.ptr(out), ctypes.c_int64(dim),
                 _native.stream_ptr(dev))
  return out.to(tdt) if tdt == torch.float16 else out


def sample_int(dim, num_results, skip=0):
  """Integer points (uint32 stored in int32) and nd of `sample`."""
  dev = _native.device()
  nd = num_digits_for(skip, num_results)
  v = device_columns(dim, nd, dev=dev)
  out = torch.empty((num_results, dim), dtype=torch.int32, device=dev)
  _native.call('tqf_sobol_points', _native.ptr(v), ctypes.c_int32(dim),
               ctypes.c_int32(nd), ctypes.c_int64(num_results),
               ctypes.c_int64(skip + 1), None, None, ctypes.c_int32(0),
               ctypes.c_int32(0), _native.ptr(out), ctypes.c_int64(dim),
               _native.stream_ptr(dev))
  return out, nd
