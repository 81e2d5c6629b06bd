"""2-D interpolation (mirror of tff.math.interpolation.interpolation_2d,
tff/math/interpolation/interpolation_2d/interpolation_2d.py:24-158).

For every x_data row a natural cubic spline in y; the row values at the
requested y are then interpolated linearly in x with constant
extrapolation.
"""
from tqf.math.interpolation import cubic
from tqf.math.interpolation import linear
from tqf.math.interpolation._util import as_tensor

__all__ = ['Interpolation2D']


class Interpolation2D:
  """x_data batch + [nx]; y_data, z_data batch + [nx, ny]."""

  def __init__(self, x_data, y_data, z_data, dtype=None, name=None):
    self._name = name or 'interpolation_2d'
    self._xdata = as_tensor(x_data, dtype)
    self._dtype = self._xdata.dtype
    self._ydata = as_tensor(y_data, self._dtype, like=self._xdata)
    self._zdata = as_tensor(z_data, self._dtype, like=self._xdata)
    self._spline_yz = cubic.build_spline(self._ydata, self._zdata)
    self._on_device = {self._xdata.device: (self._xdata, self._spline_yz)}

  def _data(self, device):
    """x_data and the y-splines on `device` (copied once per device)."""
    if device not in self._on_device:
      sp = self._spline_yz
      self._on_device[device] = (self._xdata.to(device), cubic.SplineParameters(
          sp.x_data.to(device), sp.y_data.to(device),
          sp.spline_coeffs.to(device)))
    return self._on_device[device]

  def interpolate(self, x, y, name=None):
    """Values at points (x, y), both batch + [num_points]."""
    del name
    y = as_tensor(y, self._dtype)
    x = as_tensor(x, self._dtype, like=y)
    xdata, spline = self._data(y.device)
    # batch + [nx, num_points]: every row's spline at the requested y
    xy = cubic.interpolate(y.unsqueeze(-2), spline)
    r = xy.dim()
    # [num_points] + batch + [nx]
    yx = xy.permute([r - 1] + list(range(r - 1)))
    xq = x.permute([r - 2] + list(range(r - 2))).unsqueeze(-1)
    z = linear.interpolate(xq, xdata, yx)
    return z.permute(list(range(1, r)) + [0]).squeeze(-2)
