"""Left-continuous piecewise-constant functions (tff/math/piecewise.py:19-210).

Unbatched host version: the models evaluate parameters once per grid step
on the host and hand the per-step values to the kernels.
"""
import numpy as np

__all__ = ['PiecewiseConstantFunc']


class PiecewiseConstantFunc:
  """f(x) = values[i] on (jumps[i-1], jumps[i]] (left-continuous)."""

  is_piecewise_constant = True

  def __init__(self, jump_locations, values, dtype=None, name=None):
    del name
    self._jumps = np.asarray(jump_locations, np.float64).reshape(-1)
    self._values = np.asarray(values, np.float64).reshape(-1)
    self._dtype = dtype
    if self._values.shape[0] != self._jumps.shape[0] + 1:
      raise ValueError('Event shape of `values` should have one more '
                       'element than the event shape of `jump_locations` '
                       'but are {0} and {1}'.format(self._values.shape[0],
                                                    self._jumps.shape[0]))

  def dtype(self):
    return self._dtype

  def jump_locations(self):
    return self._jumps

  def values(self):
    return self._values

  def __call__(self, x, left_continuous=True):
    side = 'left' if left_continuous else 'right'
    idx = np.searchsorted(self._jumps, np.asarray(x, np.float64), side=side)
    return self._values[idx]

  def integrate(self, x1, x2):
    """Integral of f over [x1, x2] (elementwise)."""
    x1 = np.asarray(x1, np.float64)
    x2 = np.asarray(x2, np.float64)
    knots = np.concatenate([[-np.inf], self._jumps, [np.inf]])
    lo = np.clip(x1[..., None], knots[:-1], knots[1:])
    hi = np.clip(x2[..., None], knots[:-1], knots[1:])
    return np.sum((hi - lo) * self._values, axis=-1)
