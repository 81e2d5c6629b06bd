"""Implied volatility surface (mirror of the pricing platform's
VolatilitySurface,
tff/experimental/pricing_platform/framework/market_data/volatility_surface.py
:30-176), and the calendar helpers it needs.

Dates are numpy datetime64[D] arrays (or anything `to_dates` accepts:
(year, month, day) tuples, datetime.date, datetime64, nested lists of
them).  Year fractions are actual/365 fixed, the surface's default day
count.
"""
import datetime

import numpy as np
import torch

from tqf.math.interpolation import interpolation_2d
from tqf.math.interpolation._util import as_tensor

__all__ = ['VolatilitySurface', 'dates_from_year_month_day', 'to_dates',
           'daycount_actual_365_fixed']


def dates_from_year_month_day(year, month, day):
  """datetime64[D] array from broadcast year / month / day arrays
  (tff.datetime.dates_from_year_month_day)."""
  y, m, d = np.broadcast_arrays(np.asarray(year), np.asarray(month),
                                np.asarray(day))
  out = np.empty(y.shape, 'datetime64[D]')
  for i in np.ndindex(y.shape):
    out[i] = np.datetime64(
        datetime.date(int(y[i]), int(m[i]), int(d[i])), 'D')
  return out


def _one(x):
  if isinstance(x, np.datetime64):
    return x.astype('datetime64[D]')
  if isinstance(x, datetime.date):
    return np.datetime64(x, 'D')
  if (isinstance(x, (tuple, list)) and len(x) == 3 and
      all(isinstance(v, (int, np.integer)) for v in x)):
    return np.datetime64(datetime.date(int(x[0]), int(x[1]), int(x[2])), 'D')
  return None


def to_dates(x):
  """datetime64[D] array from dates in any of the accepted forms."""
  if isinstance(x, np.ndarray) and np.issubdtype(x.dtype, np.datetime64):
    return x.astype('datetime64[D]')
  one = _one(x)
  if one is not None:
    return np.asarray(one)
  return np.asarray([to_dates(v) for v in x], 'datetime64[D]')


def daycount_actual_365_fixed(start_date, end_date, dtype=torch.float64,
                              device=None):
  """(end - start) in days / 365, broadcast (tff.datetime.
  daycount_actual_365_fixed)."""
  days = (to_dates(end_date) - to_dates(start_date)).astype(np.int64)
  return torch.as_tensor(days / 365.0, dtype=dtype, device=device)


class VolatilitySurface:
  """Implied volatilities on (expiry, strike) nodes with an interpolator
  `interpolator(expiry_times, strikes)`; default Interpolation2D (natural
  spline in strike per expiry, linear in expiry)."""

  def __init__(self, valuation_date, expiries, strikes, volatilities,
               daycount_convention=None, interpolator=None, dtype=None,
               name=None):
    if daycount_convention not in (None, 'ACTUAL_365'):
      raise NotImplementedError('only the actual/365 day count is provided')
    self._name = name or 'VolatilitySurface'
    self._dtype = dtype or torch.float64
    self._valuation_date = to_dates(valuation_date)
    self._expiries = to_dates(expiries)
    self._strikes = as_tensor(strikes, self._dtype)
    self._volatilities = as_tensor(volatilities, self._dtype)
    expiry_times = daycount_actual_365_fixed(self._valuation_date,
                                             self._expiries, self._dtype)
    if interpolator is None:
      obj = interpolation_2d.Interpolation2D(expiry_times, self._strikes,
                                             self._volatilities,
                                             dtype=self._dtype)
      self._interpolator = lambda t, x: obj.interpolate(t, x)
    else:
      self._interpolator = interpolator

  def volatility(self, strike, expiry_dates=None, expiry_times=None,
                 term=None):
    """Interpolated implied volatility at (expiry, strike)."""
    del term
    if expiry_dates is not None and expiry_times is not None:
      raise ValueError('Unexpected inputs: Both expiry_dates and expiry '
                       'times are specified')
    if expiry_times is None:
      expiries = daycount_actual_365_fixed(self._valuation_date,
                                           expiry_dates, self._dtype)
    else:
      expiries = as_tensor(expiry_times, self._dtype)
    strike = as_tensor(strike, self._dtype)
    if isinstance(expiries, torch.Tensor) and expiries.device != strike.device:
      expiries = expiries.to(strike.device)
    return self._interpolator(expiries, strike)

  def settlement_date(self):
    return self._valuation_date

  def node_expiries(self):
    return self._expiries

  def node_strikes(self):
    return self._strikes

  def node_volatilities(self):
    return self._volatilities

  def interpolator(self):
    return self._interpolator
