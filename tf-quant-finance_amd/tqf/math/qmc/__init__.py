"""Digital nets and Sobol generating matrices (mirror of tff.math.qmc).

  sobol_generating_matrices  tff/math/qmc/sobol.py:132-395
  digital_net_sample         tff/math/qmc/digital_net.py:205-419
  sobol_sample               tff/math/qmc/sobol.py:32-129
  random_digital_shift       tff/math/qmc/digital_net.py:45-95 (values drawn
                             from the engine's Philox stream, not TF's)
  random_scrambling_matrices tff/math/qmc/digital_net.py:98-152 (likewise)
  scramble_generating_matrices tff/math/qmc/digital_net.py:422-527
  tent_transform             tff/math/qmc/utils.py:124-141

Points are produced by the `tqf_sobol_points` kernel (index-0 sequence or
explicit sequence indices, XOR digital shift, optional tent transform).
"""
import ctypes
import math as _math

import numpy as np
import torch

from tqf import _native
from tqf import errors
from tqf.math.random import sobol as sobol_lib
from tqf.math.random.multivariate_normal import seed_to_key

__all__ = ['sobol_generating_matrices', 'digital_net_sample', 'sobol_sample',
           'random_digital_shift', 'random_scrambling_matrices',
           'scramble_generating_matrices', 'tent_transform']


def _log2_ceil_f32(n):
  # ceil(log(n) / log(2)) evaluated in float32 as qmc/utils.py:57-76 does
  return int(np.ceil(np.log(np.float32(n)) / np.log(np.float32(2))))


def sobol_generating_matrices(dim, num_results, num_digits,
                              validate_args=False, dtype=None, name=None):
  """[dim, ceil(log2 N)] generating matrices (host; numpy int array)."""
  del name
  log_n = _log2_ceil_f32(num_results)
  if validate_args:
    if dim <= 0 or num_results <= 0 or num_digits <= 0:
      raise errors.InvalidArgumentError(
          'dim, num_results and num_digits must be positive')
    if log_n >= 32:
      raise errors.InvalidArgumentError('log2(num_results) must be less than 32')
  v = sobol_lib.generating_columns(dim, num_digits, log_n).astype(np.int64)
  np_dtype = np.int64 if dtype in (np.int64, torch.int64, 'int64') else np.int32
  return v[:, :log_n].astype(np_dtype)


def tent_transform(value):
  return torch.where(value < 0.5, 2 * value, 2 * (1 - value))


def digital_net_sample(generating_matrices, num_results, num_digits,
                       sequence_indices=None, scrambling_matrices=None,
                       digital_shift=None, apply_tent_transform=False,
                       validate_args=False, dtype=None, name=None):
  """Samples a digital net on the GPU, shape [num_samples, dim]."""
  del name
  g = np.asarray(generating_matrices.detach().cpu() if isinstance(
      generating_matrices, torch.Tensor) else generating_matrices, np.int64)
  if scrambling_matrices is not None:
    # digital_net.py:349-353: the net of the scrambled matrices
    g = scramble_generating_matrices(g, scrambling_matrices, num_digits,
                                     validate_args=validate_args,
                                     dtype=np.int64)
  if validate_args and g.ndim != 2:
    raise errors.InvalidArgumentError('generating_matrices must have rank 2')
  dim, ncols = g.shape
  log_n = _log2_ceil_f32(num_results)
  if validate_args and log_n >= 32:
    raise errors.InvalidArgumentError('log2(num_results) must be less than 32')
  num_digits = int(num_digits)
  if num_digits > 32:
    raise ValueError('num_digits must be at most 32')
  dev = _native.device()
  cols = np.zeros((dim, 32), np.uint64)
  use = min(log_n, ncols)
  cols[:, :use] = g[:, :use].astype(np.uint64) & np.uint64(0xFFFFFFFF)
  v = torch.from_numpy(cols.astype(np.uint32).view(np.int32)).to(dev)
  seq = None
  n = int(num_results)
  if sequence_indices is not None:
    seq_np = np.asarray(sequence_indices.detach().cpu() if isinstance(
        sequence_indices, torch.Tensor) else sequence_indices, np.int64)
    if validate_args and np.any(seq_np >= num_results):
      raise errors.InvalidArgumentError(
          'values in sequence_indices must be less than num_results')
    # bits beyond log2(num_results) are ignored by the reference's loop
    seq_np = seq_np & ((1 << log_n) - 1) if log_n < 63 else seq_np
    seq = torch.as_tensor(seq_np, device=dev)
    n = seq_np.shape[0]
  shift = None
  if digital_shift is not None:
    s = np.asarray(digital_shift.detach().cpu() if isinstance(
        digital_shift, torch.Tensor) else digital_shift, np.int64)
    if validate_args and (s.ndim != 1 or s.shape[0] != dim):
      raise errors.InvalidArgumentError(
          'digital_shift must have size tf.shape(generating_matrices)[0]')
    shift = torch.from_numpy(
        (s.astype(np.uint64) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(
            np.int32)).to(dev)
  tdt = sobol_lib._torch_dtype(dtype)
  kind = 2 if tdt == torch.float64 else 1
  out = torch.empty((n, dim), dtype=torch.float64 if kind == 2 else
                    torch.float32, device=dev)
  if seq is None and sequence_indices is None:
    # indices 0..N-1 use only the low log2(N) bits (all of them)
    pass
  if n:
    _native.call('tqf_sobol_points', _native.ptr(v), ctypes.c_int32(dim),
                 ctypes.c_int32(num_digits), ctypes.c_int64(n),
                 ctypes.c_int64(0), _native.ptr(seq), _native.ptr(shift),
                 ctypes.c_int32(kind), ctypes.c_int32(int(apply_tent_transform)),
                 _native.ptr(out), ctypes.c_int64(dim),
                 _native.stream_ptr(dev))
  return out.to(tdt)


def sobol_sample(dim, num_results, sequence_indices=None, digital_shift=None,
                 scrambling_matrices=None, apply_tent_transform=False,
                 validate_args=False, dtype=None, name=None):
  """Index-0 Sobol points with nd = ceil(log2 N) digits."""
  del name
  nd = _log2_ceil_f32(num_results)
  g = sobol_generating_matrices(dim, num_results, nd,
                                validate_args=validate_args)
  return digital_net_sample(g, num_results, nd,
                            sequence_indices=sequence_indices,
                            scrambling_matrices=scrambling_matrices,
                            digital_shift=digital_shift,
                            apply_tent_transform=apply_tent_transform,
                            validate_args=validate_args, dtype=dtype)


def random_digital_shift(dim, num_digits, seed, validate_args=False,
                         dtype=None, name=None):
  """[dim] integers uniform in [2^(nd-1), 2^nd) from the engine stream.

  The reference draws them with tf.random.stateless_uniform
  (digital_net.py:155-202); that stream is not reproducible here, so the
  engine's Philox stream (sub-stream 0x5348, path 0) supplies the words.
  """
  del validate_args, name
  if num_digits < 1 or num_digits > 32:
    raise ValueError('num_digits must be in [1, 32]')
  key = seed_to_key(seed)
  words = philox_words(dim, key, stream_id=0x5348)
  lo = 1 << (num_digits - 1)
  span = lo  # maxval - minval
  vals = lo + (words.astype(np.uint64) % np.uint64(span)).astype(np.int64)
  return np.asarray(vals, np.int64 if dtype in (np.int64, torch.int64)
                    else np.int64)


def random_scrambling_matrices(dim, num_digits, seed, validate_args=False,
                               dtype=None, name=None):
  """[dim, num_digits] integers uniform in [2^(nd-1), 2^nd): the rows of
  the lower-triangular scrambling matrices, from the engine's Philox
  stream (sub-stream 0x5343); the reference draws them from TF's stateless
  uniform (digital_net.py:98-152), not reproducible here."""
  del validate_args, name
  if num_digits < 1 or num_digits > 32:
    raise ValueError('num_digits must be in [1, 32]')
  key = seed_to_key(seed)
  words = philox_words(int(dim) * int(num_digits), key, stream_id=0x5343)
  lo = 1 << (num_digits - 1)
  vals = lo + (words.astype(np.uint64) % np.uint64(lo)).astype(np.int64)
  np_dtype = np.int64 if dtype in (np.int64, torch.int64, 'int64') else np.int32
  return vals.reshape(int(dim), int(num_digits)).astype(np_dtype)


def scramble_generating_matrices(generating_matrices, scrambling_matrices,
                                 num_digits, validate_args=False, dtype=None,
                                 name=None):
  """Linear matrix scrambling (digital_net.py:422-527): entry (d, j) of the
  result is the XOR over s < num_digits of (S[d, s] >> s) where bit
  num_digits - 1 - s of G[d, j] is set.  Host integers (the matrices are
  [dim, log2 N])."""
  del name
  g = np.asarray(generating_matrices.detach().cpu() if isinstance(
      generating_matrices, torch.Tensor) else generating_matrices, np.int64)
  sm = np.asarray(scrambling_matrices.detach().cpu() if isinstance(
      scrambling_matrices, torch.Tensor) else scrambling_matrices, np.int64)
  nd = int(num_digits)
  if validate_args:
    if g.ndim != sm.ndim:
      raise errors.InvalidArgumentError(
          'input matrices must have the same rank')
    if nd <= 0:
      raise errors.InvalidArgumentError('num_digits must be positive')
  out = np.zeros_like(g)
  for s in range(nd):
    col = (sm[:, s:s + 1] >> s)  # [dim, 1]
    bit = (g >> (nd - 1 - s)) & 1
    out ^= np.where(bit == 1, col, 0)
  if dtype is None:
    np_dtype = (generating_matrices.dtype if isinstance(
        generating_matrices, np.ndarray) else np.int32)
  else:
    np_dtype = np.int64 if dtype in (np.int64, torch.int64, 'int64') \
        else np.int32
  return out.astype(np_dtype)


def philox_words(n, key, stream_id=0):
  """n raw uint32 Philox words (path 0, counters 0..), computed on host.

  Used for the handful of set-up constants (digital shifts); the host
  routine follows the same Philox4x32-10 definition as the kernels.
  """
  m0, m1, w0, w1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
  out = []
  b = 0
  while len(out) < n:
    c = [0, b, stream_id, 0]
    k0, k1 = key & 0xFFFFFFFF, (key >> 32) & 0xFFFFFFFF
    for _ in range(10):
      p0, p1 = m0 * c[0], m1 * c[2]
      c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
           ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
      k0 = (k0 + w0) & 0xFFFFFFFF
      k1 = (k1 + w1) & 0xFFFFFFFF
    out.extend(c)
    b += 1
  return np.asarray(out[:n], np.uint64)
