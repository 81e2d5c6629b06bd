"""Piecewise-constant functions (mirror of tff.math.piecewise,
tff/math/piecewise.py:19-340).

  PiecewiseConstantFunc   piecewise.py:19-210: jump locations
                          batch_shape + [J], values batch_shape + [J + 1] +
                          event_shape; evaluation (left- or right-continuous)
                          and integration between end points
  find_interval_index     piecewise.py:213-273

Host (numpy) code: the models evaluate their parameters once per grid
step on the host and hand the per-step values to the kernels, so nothing
here runs per path.  Results are numpy arrays in the function's dtype.
"""
import numpy as np

__all__ = ['PiecewiseConstantFunc', 'find_interval_index']


def _np_dtype(dtype, default):
  if dtype is None:
    return default
  name = getattr(dtype, 'name', None) or str(dtype)
  if 'float64' in name or 'double' in name:
    return np.float64
  if 'float32' in name or name == 'float':
    return np.float32
  return np.dtype(dtype).type


def _array(x, dtype=None):
  try:
    import torch
    if isinstance(x, torch.Tensor):
      x = x.detach().cpu().numpy()
  except ImportError:
    pass
  return np.asarray(x, dtype=dtype)


class PiecewiseConstantFunc:
  """f(x) = values[..., i] on the i-th interval between jump locations."""

  is_piecewise_constant = True

  def __init__(self, jump_locations, values, dtype=None, name=None):
    self._name = name or 'PiecewiseConstantFunc'
    jl = _array(jump_locations)
    self._dtype = _np_dtype(dtype, jl.dtype.type if jl.dtype.kind == 'f'
                            else np.float32)
    self._jumps = _array(jump_locations, self._dtype)
    self._values = _array(values, self._dtype)
    if self._jumps.ndim == 0:
      self._jumps = self._jumps.reshape(1)
    self._batch_rank = self._jumps.ndim - 1
    br = self._batch_rank
    if self._values.shape[:br] != self._jumps.shape[:-1]:
      raise ValueError(
          'Batch shapes of `values` and `jump_locations` should be the same '
          'but are {0} and {1}'.format(self._values.shape[:-1],
                                       self._jumps.shape[:-1]))
    if (self._values.ndim <= br or
        self._values.shape[br] - 1 != self._jumps.shape[-1]):
      raise ValueError('Event shape of `values` should have one more '
                       'element than the event shape of `jump_locations` '
                       'but are {0} and {1}'.format(self._values.shape[-1],
                                                    self._jumps.shape[-1]))

  def dtype(self):
    return self._dtype

  def name(self):
    return self._name

  def jump_locations(self):
    return self._jumps

  def values(self):
    return self._values

  def _batched(self, x):
    """x as [B, P] against jumps [B, J] and values [B, J + 1, event]."""
    br = self._batch_rank
    batch = self._jumps.shape[:-1]
    x = _array(x, self._dtype)
    no_batch = x.ndim <= br or x.shape[:br] != batch
    if no_batch and br:
      # a batch-free x is evaluated for every batch member
      x = np.broadcast_to(x, batch + x.shape[-1:] if x.ndim else batch + (1,))
    pts_shape = x.shape[br:]
    b = int(np.prod(batch)) if br else 1
    xs = x.reshape(b, -1)
    jumps = self._jumps.reshape(b, -1)
    vals = self._values.reshape((b,) + self._values.shape[br:])
    return xs, jumps, vals, batch + pts_shape

  def __call__(self, x, left_continuous=True, name=None):
    """Values at x (batch_shape + [num_points]) -> batch_shape +
    [num_points] + event_shape (piecewise.py:129-163)."""
    del name
    side = 'left' if left_continuous else 'right'
    xs, jumps, vals, out_shape = self._batched(x)
    res = np.stack([vals[i][np.searchsorted(jumps[i], xs[i], side=side)]
                    for i in range(xs.shape[0])])
    return res.reshape(out_shape + vals.shape[2:]).astype(self._dtype)

  def integrate(self, x1, x2, name=None):
    """Integral over [x1, x2] (batch_shape + [num_points]) -> batch_shape +
    [num_points] + event_shape (piecewise.py:165-210)."""
    del name
    a, jumps, vals, out_shape = self._batched(x1)
    b2, _, _, _ = self._batched(x2)
    out = []
    for i in range(a.shape[0]):
      knots = np.concatenate([[-np.inf], jumps[i], [np.inf]])
      lo = np.clip(a[i][:, None], knots[:-1], knots[1:])
      hi = np.clip(b2[i][:, None], knots[:-1], knots[1:])
      w = hi - lo  # [P, J + 1]
      out.append(np.tensordot(w, vals[i], axes=([1], [0])))
    res = np.stack(out)
    return res.reshape(out_shape + vals.shape[2:]).astype(self._dtype)


def find_interval_index(query_xs, interval_lower_xs,
                        last_interval_is_closed=False, dtype=None, name=None):
  """Index of the half-open interval [x_i, x_{i+1}) holding each query
  point: -1 before x_0, n for x >= x_n; with last_interval_is_closed the
  last interval includes its right end (piecewise.py:213-273).  int32."""
  del name
  q = _array(query_xs, dtype)
  lows = _array(interval_lower_xs, q.dtype)
  idx = np.searchsorted(lows, q, side='right').astype(np.int32) - 1
  if last_interval_is_closed:
    last = lows.shape[-1] - 1
    cap = last - ((idx == last) & (q <= lows[last])).astype(np.int32)
    idx = np.minimum(idx, cap)
  return idx.astype(np.int32)
