"""Piecewise-linear interpolation (mirror of
tff.math.interpolation.linear.interpolate,
tff/math/interpolation/linear/linear_interpolation.py:27-196).

Knot index by searchsorted(side='left'), so a point on a knot takes the
interval to its left; outside the knots the value is extended with
`left_slope` / `right_slope` (0: constant); repeated knots give a zero
slope.  Batch dimensions broadcast.
"""
import torch

from tqf.math.interpolation._util import as_tensor, broadcast_batch
from tqf.math.interpolation._util import divide_no_nan

__all__ = ['interpolate']


def interpolate(x, x_data, y_data, left_slope=None, right_slope=None,
                validate_args=False, optimize_for_tpu=False, dtype=None,
                name=None):
  del optimize_for_tpu, name
  x = as_tensor(x, dtype)
  x_data = as_tensor(x_data, x.dtype, like=x)
  y_data = as_tensor(y_data, x.dtype, like=x)
  scalar = x.dim() == 1 and x_data.dim() == 1 and y_data.dim() == 1
  if validate_args:
    if bool((x_data[..., 1:] < x_data[..., :-1]).any()):
      raise ValueError('x_data is not sorted in non-decreasing order.')
    if x_data.shape != y_data.shape:
      raise ValueError('x_data and y_data must have the same shape')
  x, x_data, y_data = broadcast_batch(x, x_data, y_data)
  ls = (torch.zeros((), dtype=x.dtype, device=x.device) if left_slope is None
        else as_tensor(left_slope, x.dtype, like=x))
  rs = (torch.zeros((), dtype=x.dtype, device=x.device) if right_slope is None
        else as_tensor(right_slope, x.dtype, like=x))
  n = x_data.shape[-1]
  upper = torch.searchsorted(x_data.contiguous(), x.detach().contiguous(),
                             right=False)
  at_min = upper == 0
  at_max = upper == n
  values_min = y_data[..., :1] + ls * (x - x_data[..., :1])
  values_max = y_data[..., -1:] + rs * (x - x_data[..., -1:])
  lo = (upper - 1).clamp(min=0)
  hi = upper.clamp(max=n - 1)
  xl, xu = x_data.gather(-1, lo), x_data.gather(-1, hi)
  yl, yu = y_data.gather(-1, lo), y_data.gather(-1, hi)
  diff = xu - xl
  diff = torch.where(at_min | at_max, diff + 1, diff)
  out = yl + (x - xl) * divide_no_nan(yu - yl, diff)
  out = torch.where(at_min, values_min, out)
  out = torch.where(at_max, values_max, out)
  return out.reshape(-1) if scalar else out
