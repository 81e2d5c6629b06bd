"""Oracle: Halton sequences.

  * `sample`  tff/math/random_ops/halton/halton_impl.py:59-285: element k of
    the sequence is the radical inverse of the integer k + 1 (1-based,
    :386-408) in the prime base of each axis, truncated to the number of
    base-b digits of the dtype's largest index (2^24 - 1 for float32,
    2^53 - 1 for float64, :287-305 and :411-429);
  * randomized (Owen 2017, :252-285 and :288-330): the digit of position j
    on axis d goes through permutation perms[j, offset_d + digit] (digits
    past the index's length are the zeros that get permuted too, up to the
    truncation length), then a zero-correction u_d / b_d^max_size is added.

The permutations and u_d are inputs (the reference draws them from TF's
stateless RNG, which cannot run here: those values are parity unpinned).
Pinned by: the reference's known values (halton_test.py:30-37: the first
five points in bases 2 and 3), sequence_indices = range(n) reproducing
num_results = n (:50-61), and the radical inverse computed independently
with Python fractions on small indices (tests/test_oracle.py).
"""
import numpy as np

MAX_INDEX = {np.float32: 2**24 - 1, np.float64: 2**53 - 1}
NUM_COEFFS = {np.float32: 24, np.float64: 54}


def primes(n):
  """The first n primes (sieve)."""
  if n < 1:
    return np.zeros(0, np.int64)
  limit = 16
  while True:
    sieve = np.ones(limit + 1, bool)
    sieve[:2] = False
    for p in range(2, int(limit**0.5) + 1):
      if sieve[p]:
        sieve[p * p::p] = False
    ps = np.nonzero(sieve)[0]
    if ps.shape[0] >= n:
      return ps[:n].astype(np.int64)
    limit *= 2


def max_digits(bases, dtype):
  """floor(log(max_index) / log(b)) + 1, evaluated in float64 (the
  reference computes it in numpy from Python ints, halton_impl.py:530-534).
  """
  m = float(MAX_INDEX[dtype])
  return (np.floor(np.log(m) / np.log(np.asarray(bases, np.float64))) +
          1).astype(np.int64)


def sample(dim, num_results=None, sequence_indices=None, perms=None,
           zero_correction=None, dtype=np.float32):
  """[n, dim] Halton points; perms [num_coeffs, sum(bases)] and
  zero_correction [dim] (already divided by b^max_size) randomize."""
  if sequence_indices is None:
    sequence_indices = np.arange(num_results)
  idx = np.asarray(sequence_indices, np.int64) + 1
  bases = primes(dim)
  sizes = max_digits(bases, dtype)
  offsets = np.concatenate([[0], np.cumsum(bases)[:-1]])
  radix_sum = int(bases.sum())
  out = np.zeros((idx.shape[0], dim), np.float64)
  for d in range(dim):
    b = int(bases[d])
    rem = idx.copy()
    acc = np.zeros(idx.shape[0], np.float64)
    w = 1.0
    for j in range(int(sizes[d])):
      c = rem % b
      rem //= b
      if perms is not None:
        c = np.asarray(perms).reshape(-1)[j * radix_sum + offsets[d] + c]
      w *= b
      acc += c / w
    if zero_correction is not None:
      acc += zero_correction[d]
    out[:, d] = acc
  return out.astype(dtype)
