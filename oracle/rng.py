"""Oracle: the engine's counter-based pseudo-random normal stream.

The reference's PSEUDO / STATELESS streams come from TensorFlow's Philox
kernels (tff/math/random_ops/multivariate_normal.py:248-273), whose values
cannot be reproduced without TF: parity with those streams is unpinned and
the engine defines its own stream, restated here.

  * Philox4x32-10, J. Salmon, M. Moraes, R. Dror, D. Shaw, "Parallel random
    numbers: as easy as 1, 2, 3", SC'11.  Pinned by the three published
    known-answer vectors of that paper's reference implementation
    (tests/golden/philox_kat.json).
  * Box-Muller on each (x0, x1) and (x2, x3) pair.  Pinned by moment and
    distribution tests (mean 0, variance 1, KS against N(0,1)) and by the
    closed form r = sqrt(-2 ln u) checked on hand-picked counters.

Stream layout (shared with the HIP kernels): normal number j of path p in
sub-stream s is word (j % 4) of Philox(counter=(p_lo, j // 4, s, p_hi),
key=(seed_lo, seed_hi)).
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint64(0x9E3779B9)
W1 = np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(counter, key):
  """counter: uint64 array [..., 4] (values < 2^32); key: [..., 2]."""
  c = [np.asarray(counter[..., i], dtype=np.uint64) for i in range(4)]
  k0 = np.asarray(key[..., 0], dtype=np.uint64) & _MASK
  k1 = np.asarray(key[..., 1], dtype=np.uint64) & _MASK
  for _ in range(10):
    p0 = M0 * c[0]
    p1 = M1 * c[2]
    hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
    hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
    c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    k0 = (k0 + W0) & _MASK
    k1 = (k1 + W1) & _MASK
  return np.stack(c, axis=-1)


def seed_key(seed):
  """Python int seed -> (lo, hi) words; a [2] sequence is taken as is."""
  if np.ndim(seed) == 1:
    return int(seed[0]) & 0xFFFFFFFF, int(seed[1]) & 0xFFFFFFFF
  seed = int(seed) & 0xFFFFFFFFFFFFFFFF
  return seed & 0xFFFFFFFF, seed >> 32


def raw_words(num_paths, num_draws, seed, stream=0, path_offset=0):
  """uint64 [num_paths, num_draws] Philox words of the stream layout."""
  lo, hi = seed_key(seed)
  nb = (num_draws + 3) // 4
  p = np.arange(num_paths, dtype=np.uint64) + np.uint64(path_offset)
  b = np.arange(nb, dtype=np.uint64)
  ctr = np.zeros((num_paths, nb, 4), dtype=np.uint64)
  ctr[..., 0] = (p & _MASK)[:, None]
  ctr[..., 1] = b[None, :]
  ctr[..., 2] = np.uint64(stream)
  ctr[..., 3] = (p >> np.uint64(32))[:, None]
  key = np.zeros((num_paths, nb, 2), dtype=np.uint64)
  key[..., 0] = lo
  key[..., 1] = hi
  w = philox4x32_10(ctr, key).reshape(num_paths, nb * 4)
  return w[:, :num_draws]


def box_muller(words, dtype):
  """Pairs (w0, w1) -> (r cos 2πt, r sin 2πt), r = sqrt(-2 ln u).

  float64: u = (w0 + 1) 2^-32 in (0, 1], t = w1 2^-32.
  float32: u = ((w0 >> 8) + 1) 2^-24, t = (w1 >> 8) 2^-24 (24-bit uniforms).
  Evaluated in float64 here; the float32 kernel is compared with a tolerance.
  """
  w = np.asarray(words, dtype=np.uint64)
  n = w.shape[-1]
  pad = (-n) % 2
  if pad:
    w = np.concatenate([w, np.zeros(w.shape[:-1] + (1,), np.uint64)], -1)
  a, b = w[..., 0::2], w[..., 1::2]
  if dtype == np.float32:
    u = ((a >> np.uint64(8)).astype(np.float64) + 1.0) * 2.0**-24
    t = (b >> np.uint64(8)).astype(np.float64) * 2.0**-24
  else:
    u = (a.astype(np.float64) + 1.0) * 2.0**-32
    t = b.astype(np.float64) * 2.0**-32
  r = np.sqrt(-2.0 * np.log(u))
  z = np.empty(a.shape[:-1] + (a.shape[-1] * 2,), dtype=np.float64)
  z[..., 0::2] = r * np.cos(2 * np.pi * t)
  z[..., 1::2] = r * np.sin(2 * np.pi * t)
  return z[..., :n]


def normals(num_paths, num_draws, seed, stream=0, dtype=np.float64,
            antithetic=False):
  """[num_paths, num_draws] normals of the engine's PSEUDO stream.

  Antithetic: rows i and i + N/2 are (z, -z), the layout of
  multivariate_normal.py:276-311 (the first N/2 rows are drawn).
  """
  if antithetic:
    if num_paths % 2:
      raise ValueError('antithetic sampling needs an even number of paths')
    half = normals(num_paths // 2, num_draws, seed, stream, dtype)
    return np.concatenate([half, -half], axis=0)
  # pairs are (w0, w1), (w2, w3) of each 4-word block, also for the last
  # partial block
  nb4 = (num_draws + 3) // 4 * 4
  w = raw_words(num_paths, nb4, seed, stream)
  return box_muller(w, dtype)[:, :num_draws]
