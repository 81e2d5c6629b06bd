"""Oracle: closed-form prices and Greeks.

  * Black-Scholes vanilla price, tff/black_scholes/vanilla_prices.py:28-183
    (forward form, e^{-rT}[F N(d1) - K N(d2)]), restated with scipy.special.
  * delta, vega, rho: the textbook derivatives of that formula.

Pinned by: put-call parity, intrinsic-value limits (sigma -> 0), and the
notebook's published Greek values (Monte_Carlo_Euler_Scheme.ipynb:876-879
uses this formula as its truth).
"""
import numpy as np
from scipy import special


def _ncdf(x):
  return 0.5 * special.erfc(-x / np.sqrt(2.0))


def bs_price(s0, k, t, r, sigma, is_call=True, q=0.0):
  s0, k, t, sigma = (np.asarray(a, np.float64) for a in (s0, k, t, sigma))
  f = s0 * np.exp((r - q) * t)
  sd = sigma * np.sqrt(t)
  d1 = np.log(f / k) / sd + sd / 2
  d2 = d1 - sd
  df = np.exp(-r * t)
  call = df * (f * _ncdf(d1) - k * _ncdf(d2))
  if is_call:
    return call
  return call - df * (f - k)


def bs_delta(s0, k, t, r, sigma, q=0.0):
  sd = sigma * np.sqrt(t)
  d1 = (np.log(s0 / k) + (r - q) * t) / sd + sd / 2
  return np.exp(-q * t) * _ncdf(d1)


def bs_vega(s0, k, t, r, sigma, q=0.0):
  sd = sigma * np.sqrt(t)
  d1 = (np.log(s0 / k) + (r - q) * t) / sd + sd / 2
  return s0 * np.exp(-q * t) * np.exp(-0.5 * d1 * d1) / np.sqrt(
      2 * np.pi) * np.sqrt(t)


def bs_rho(s0, k, t, r, sigma, q=0.0):
  sd = sigma * np.sqrt(t)
  d1 = (np.log(s0 / k) + (r - q) * t) / sd + sd / 2
  d2 = d1 - sd
  return k * t * np.exp(-r * t) * _ncdf(d2)
