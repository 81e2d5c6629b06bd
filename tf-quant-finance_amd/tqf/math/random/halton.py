"""Halton sequences on the GPU (mirror of tff.math.random.halton).

  sample        tff/math/random_ops/halton/halton_impl.py:59-285
  HaltonParams  halton_impl.py:42-56

Element k is the radical inverse of k + 1 in the k-th prime base per axis,
truncated to the number of base-b digits of the dtype's largest index
(2^24 - 1 for float32, 2^53 - 1 for float64; the count is computed in
float64 as the reference does).  `randomized=True` applies Owen's (2017)
digit-permutation scrambling: one uniform permutation per (digit position,
axis) and a zero correction u_d / b_d^digits.  The reference draws them with
TF's stateless RNG, which is not reproducible here: the engine draws them
from numpy's Philox generator keyed by `seed`, and returns them in
`HaltonParams` so a run can be repeated or extended with
`randomization_params`.  Points come from the `tqf_halton_points` kernel.
"""
import ctypes
import dataclasses
import random as _pyrandom

import numpy as np
import torch

from tqf import _native
from tqf import errors

__all__ = ['sample', 'HaltonParams']

_MAX_DIMENSION = 1000
_MAX_INDEX = {torch.float32: 2**24 - 1, torch.float64: 2**53 - 1}
_NUM_COEFFS = {torch.float32: 24, torch.float64: 54}
_primes_cache = []


@dataclasses.dataclass
class HaltonParams:
  """Randomisation of a Halton sequence: perms [num_coeffs, sum(bases)]
  (int32 device tensor) and zero_correction [dim] (device tensor)."""
  perms: torch.Tensor
  zero_correction: torch.Tensor

  def __iter__(self):
    return iter((self.perms, self.zero_correction))


def primes(n):
  """The first n primes (sieve of Eratosthenes)."""
  if len(_primes_cache) < n:
    limit = 16
    while True:
      sieve = np.ones(limit + 1, bool)
      sieve[:2] = False
      for p in range(2, int(limit**0.5) + 1):
        if sieve[p]:
          sieve[p * p::p] = False
      ps = np.nonzero(sieve)[0]
      if ps.shape[0] >= n:
        _primes_cache[:] = list(ps)
        break
      limit *= 2
  return np.asarray(_primes_cache[:n], np.int64)


def _max_digits(bases, dtype):
  m = float(_MAX_INDEX[dtype])
  return (np.floor(np.log(m) / np.log(bases.astype(np.float64))) + 1).astype(
      np.int64)


def _torch_dtype(dtype):
  if dtype is None:
    return torch.float32
  if isinstance(dtype, torch.dtype):
    return dtype
  return torch.float64 if np.dtype(dtype) == np.float64 else torch.float32


def _randomization(dim, bases, sizes, seed, dtype, dev):
  key = _pyrandom.getrandbits(63) if seed is None else int(
      np.asarray(seed).reshape(-1)[0]) & 0x7FFFFFFFFFFFFFFF
  gen = np.random.Generator(np.random.Philox(key=key))
  nc = _NUM_COEFFS[dtype]
  radix_sum = int(bases.sum())
  # one uniform permutation per (digit position, axis): sort random keys
  # within each base-sized segment of every row
  seg = np.repeat(np.arange(dim), bases)
  keys = gen.random((nc, radix_sum))
  order = np.lexsort((keys, np.broadcast_to(seg, keys.shape)), axis=-1)
  offsets = np.concatenate([[0], np.cumsum(bases)[:-1]])
  perms = (order - offsets[seg][None, :]).astype(np.int32)
  zc = gen.random(dim) / bases.astype(np.float64)**sizes
  return HaltonParams(torch.as_tensor(perms, device=dev),
                      torch.as_tensor(zc, dtype=torch.float64, device=dev))


def _launch(dim, n, first, seq, params, dtype, normal, dev):
  bases = primes(dim)
  sizes = _max_digits(bases, dtype)
  offsets = np.concatenate([[0], np.cumsum(bases)[:-1]]).astype(np.int32)
  tabs = torch.as_tensor(np.stack([bases, sizes, offsets]).astype(np.int32),
                         device=dev)
  out = torch.empty((n, dim), dtype=dtype, device=dev)
  kind = (1 if dtype == torch.float32 else 2) + (2 if normal else 0)
  perms = zc = None
  if params is not None:
    perms = params.perms.to(device=dev, dtype=torch.int32).contiguous()
    zc = params.zero_correction.to(device=dev,
                                   dtype=torch.float64).contiguous()
  if n:
    _native.call('tqf_halton_points', ctypes.c_void_p(tabs[0].data_ptr()),
                 ctypes.c_void_p(tabs[1].data_ptr()),
                 ctypes.c_void_p(tabs[2].data_ptr()), ctypes.c_int32(dim),
                 ctypes.c_int64(n), ctypes.c_int64(first), _native.ptr(seq),
                 _native.ptr(perms), ctypes.c_int32(int(bases.sum())),
                 _native.ptr(zc), ctypes.c_int32(kind), _native.ptr(out),
                 ctypes.c_int64(dim), _native.stream_ptr(dev))
  return out


def sample(dim, num_results=None, sequence_indices=None, randomized=True,
           randomization_params=None, seed=None, validate_args=False,
           dtype=None, name=None, _normal=False):
  """Returns (points [n, dim], HaltonParams or None) as the reference."""
  del name
  if (num_results is None) == (sequence_indices is None):
    raise ValueError('Either `num_results` or `sequence_indices` must be'
                     ' specified but not both.')
  dtype = _torch_dtype(dtype)
  dev = _native.device()
  dim = int(dim)
  seq = None
  if sequence_indices is not None:
    seq_np = np.asarray(sequence_indices.detach().cpu() if isinstance(
        sequence_indices, torch.Tensor) else sequence_indices,
                        np.int64).reshape(-1)
    n, first = seq_np.shape[0], 0
    max_idx = int(seq_np.max()) if n else 0
    seq = torch.as_tensor(seq_np, device=dev)
  else:
    n, first = int(num_results), 0
    max_idx = n - 1
  if validate_args:
    if max_idx + 1 > _MAX_INDEX[dtype]:
      raise errors.InvalidArgumentError(
          'Maximum sequence index exceeded. Maximum index for dtype %s is %d.'
          % (dtype, _MAX_INDEX[dtype]))
    if dim < 1:
      raise errors.InvalidArgumentError('`dim` should be greater than 1')
    if dim > _MAX_DIMENSION:
      raise errors.InvalidArgumentError(
          '`dim` should be less or equal than %d' % _MAX_DIMENSION)
  if dim < 1 or dim > _MAX_DIMENSION:
    raise ValueError('`dim` must be in [1, %d]' % _MAX_DIMENSION)
  params = None
  if randomized:
    params = randomization_params
    if params is None:
      bases = primes(dim)
      params = _randomization(dim, bases, _max_digits(bases, dtype), seed,
                              dtype, dev)
    elif not isinstance(params, HaltonParams):
      params = HaltonParams(*params)
  out = _launch(dim, n, first, seq, params, dtype, _normal, dev)
  return out, params
