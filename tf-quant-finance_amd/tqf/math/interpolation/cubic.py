"""Cubic spline interpolation (mirror of tff.math.interpolation.cubic).

  build_spline   tff/math/interpolation/cubic/cubic_interpolation.py:52-160
  interpolate    :163-253
  coefficients   :256-359 (natural), :362-431 (clamped / fixed first
                 derivative), :434-518 (system assembly)

The stored coefficients are half the spline's second derivatives at the
knots (the tridiagonal system has 3 x slope differences on the right), and
interpolation on knot interval [x0, x1] with t = (x - x0) / (x1 - x0) is
    y0 + t (y1 - y0) + dx^2 t (t - 1) ((t + 1) 2 c1 - (t - 2) 2 c0) / 6
(Sedgewick, Algorithms in C, 1990, pp. 545-550).  Outside the knots the
value is held constant.  Repeated knots are allowed at the right end
(they pad splines of different lengths into one batch); rows of the system
that touch a zero-length interval are replaced by the identity, as the
reference does.

Everything is written so that torch autograd and forward-mode AD
(torch.func.jvp) give the piecewise-polynomial derivative inside the knots
and 0 outside, with no NaN from the masked branches.
"""
import dataclasses
import enum

import torch

from tqf.math.interpolation._util import as_tensor, broadcast_batch

__all__ = ['BoundaryConditionType', 'SplineParameters', 'build',
           'build_spline', 'interpolate']


class BoundaryConditionType(enum.Enum):
  """Natural (zero second derivative), clamped (zero first derivative) or a
  fixed first derivative at both ends."""
  NATURAL = 1
  CLAMPED = 2
  FIXED_FIRST_DERIVATIVE = 3


@dataclasses.dataclass
class SplineParameters:
  x_data: torch.Tensor
  y_data: torch.Tensor
  spline_coeffs: torch.Tensor


def _tridiagonal_solve(sup, diag, sub, rhs):
  """Thomas algorithm without pivoting over the last dimension (sup[i] is
  the (i, i+1) entry, sub[i] the (i, i-1) entry), batched over the rest."""
  n = diag.shape[-1]
  if n == 0:
    return rhs
  cp = [None] * n
  dp = [None] * n
  cp[0] = sup[..., 0] / diag[..., 0]
  dp[0] = rhs[..., 0] / diag[..., 0]
  for i in range(1, n):
    m = diag[..., i] - sub[..., i] * cp[i - 1]
    cp[i] = sup[..., i] / m
    dp[i] = (rhs[..., i] - sub[..., i] * dp[i - 1]) / m
  out = [None] * n
  out[n - 1] = dp[n - 1]
  for i in range(n - 2, -1, -1):
    out[i] = dp[i] - cp[i] * out[i + 1]
  return torch.stack(out, -1)


def _coefficients(x, y, kind, left, right):
  dx = x[..., 1:] - x[..., :-1]
  dx_safe = torch.where(dx == 0, torch.ones_like(dx), dx)
  dd = torch.where(dx == 0, torch.zeros_like(dx),
                   (y[..., 1:] - y[..., :-1]) / dx_safe)
  rhs = -3.0 * (dd[..., :-1] - dd[..., 1:])
  diag = 2.0 * (x[..., 2:] - x[..., :-2])
  sup = dx[..., 1:]
  sub = dx[..., :-1]
  zero = torch.zeros_like(dx[..., :1])
  one = torch.ones_like(dx[..., :1])
  if kind == BoundaryConditionType.NATURAL:
    corr = (sup == 0) | (sub == 0)
    diag = torch.where(corr, torch.ones_like(diag), diag)
    sup_c = torch.where(sub == 0, torch.zeros_like(sup), sup)
    sub_c = torch.where(sup == 0, torch.zeros_like(sub), sub)
    rhs = torch.where(corr, torch.zeros_like(rhs), rhs)
    inner = _tridiagonal_solve(sup_c, diag, sub_c, rhs)
    return torch.cat([zero, inner, zero], -1)
  # clamped / fixed first derivative: the full n x n system
  diag = torch.cat([2.0 * dx[..., :1], diag, zero], -1)
  sup = torch.cat([dx[..., :1], sup, zero], -1)
  sub = torch.cat([zero, sub, dx[..., -1:]], -1)
  dxx = torch.cat([one, dx, zero], -1)
  dx_right, dx_left = dxx[..., 1:], dxx[..., :-1]
  right_boundary = (dx_right == 0) & (dx_left != 0)
  diag = torch.where(right_boundary, 2.0 * dx_left, diag)
  diag = torch.where(dx_left == 0, torch.ones_like(diag), diag)
  if kind == BoundaryConditionType.FIXED_FIRST_DERIVATIVE:
    lb = as_tensor(left, like=x)
    rb = as_tensor(right, like=x)
    lb = lb.reshape(1) if lb.dim() == 0 else lb
    rb = rb.reshape(1) if rb.dim() == 0 else rb
  else:
    lb = rb = zero
  top = 3.0 * (dd[..., :1] - lb[..., :1])
  rhs = torch.cat([top.expand_as(zero), rhs, zero], -1)
  dd_left = torch.cat([one, dd], -1)
  bottom = -3.0 * (dd_left - rb[..., :1])
  rhs = torch.where(right_boundary, bottom, rhs)
  rhs = torch.where(dd_left == 0, torch.zeros_like(rhs), rhs)
  return _tridiagonal_solve(sup, diag, sub, rhs)


def build(x_data, y_data, boundary_condition_type=None,
          left_boundary_value=None, right_boundary_value=None,
          validate_args=False, dtype=None, name=None):
  """Spline through (x_data, y_data) along the last dimension; x_data
  non-decreasing, batch dimensions broadcast."""
  del name
  kind = boundary_condition_type or BoundaryConditionType.NATURAL
  x = as_tensor(x_data, dtype)
  y = as_tensor(y_data, dtype or x.dtype, like=x)
  if validate_args and bool((x[..., 1:] < x[..., :-1]).any()):
    raise ValueError('x_data is not sorted in non-decreasing order.')
  x, y = broadcast_batch(x, y)
  if kind == BoundaryConditionType.FIXED_FIRST_DERIVATIVE and (
      left_boundary_value is None or right_boundary_value is None):
    raise ValueError(
        'Expected non-empty left_boundary_value/right_boundary_value when '
        'boundary_condition_type is FIXED_FIRST_DERIVATIVE, actual '
        'left_boundary_value {0}, actual right_boundary_value {1}'.format(
            left_boundary_value, right_boundary_value))
  coeffs = _coefficients(x, y, kind, left_boundary_value,
                         right_boundary_value)
  return SplineParameters(x_data=x, y_data=y, spline_coeffs=coeffs)


build_spline = build


def interpolate(x, spline_data, optimize_for_tpu=False, dtype=None,
                name=None):
  """Spline values at x (batch shape broadcast with the spline's);
  constant outside [min x_data, max x_data]."""
  del optimize_for_tpu, name
  xd = spline_data.x_data
  x = as_tensor(x, dtype or xd.dtype)
  dev = x.device
  x, xd, yd, c = broadcast_batch(x, xd.to(dev), spline_data.y_data.to(dev),
                                 spline_data.spline_coeffs.to(dev))
  n = xd.shape[-1]
  idx = torch.searchsorted(xd.contiguous(), x.detach().contiguous(),
                           right=True) - 1
  lo = idx.clamp(min=0)
  hi = (idx + 1).clamp(max=n - 1)
  x0, x1 = xd.gather(-1, lo), xd.gather(-1, hi)
  y0, y1 = yd.gather(-1, lo), yd.gather(-1, hi)
  c0, c1 = c.gather(-1, lo), c.gather(-1, hi)
  dx = x1 - x0
  pos = dx > 0
  t = (x - x0) / torch.where(pos, dx, torch.ones_like(dx))
  t = torch.where(pos, t, torch.zeros_like(t))
  df = ((t + 1.0) * c1 * 2.0) - ((t - 2.0) * c0 * 2.0)
  df1 = df * t * (t - 1) / 6.0
  result = y0 + t * (y1 - y0) + dx * dx * df1
  upper = xd.max(-1, keepdim=True).values
  lower = xd.min(-1, keepdim=True).values
  inside = (x <= upper) & (x >= lower)
  return torch.where(inside, result, torch.where(x > upper, y0, y1))
