"""Oracle: local volatility model with a tabulated surface.

  * drift / volatility of LocalVolatilityModel in ln S coordinates,
    tff/experimental/local_volatility/local_volatility_model.py:461-476:
    drift = r - q - sigma(t, S)^2 / 2, vol = sigma(t, S) (dim 1).
  * `table_sigma`: the engine's surface format -- sigma on times t_i and a
    uniform ln S axis, bilinear in (t, ln S), flat outside (or, with
    `outside`, given values below / above the nodes); t inside
    [t_i, t_{i+1}) interpolates rows i, i+1, t below t_0 uses row 0.

Sampled with oracle.euler (the reference's Euler step).  Pinned by: a flat
surface sigma = s0 reduces the model to log-Euler GBM, exact in law (prices
within 3 SE of Black-Scholes, tests/test_gpu_core.py); a surface linear in
ln S is reproduced exactly by the bilinear lookup (tests/test_oracle.py).
"""
import numpy as np


def table_sigma(times, x0, dx, vols, t, x, outside=None):
  """outside [nt, 2]: values below the first and above the last node (used
  as they are, no interpolation in ln S); None: flat extrapolation."""
  vols = np.asarray(vols, np.float64)
  times = np.asarray(times, np.float64)
  nt, nx = vols.shape
  fraw = (np.asarray(x, np.float64) - x0) / dx
  fx = np.clip(fraw, 0, nx - 1)
  j = np.minimum(np.floor(fx).astype(np.int64), nx - 2)
  wx = fx - j
  row = 0
  while row + 1 < nt and times[row + 1] <= t:
    row += 1
  r1 = row + 1 if row + 1 < nt else row
  wt = 0.0
  if row + 1 < nt and t > times[row]:
    wt = (t - times[row]) / (times[row + 1] - times[row])
  s0 = vols[row, j] + wx * (vols[row, j + 1] - vols[row, j])
  s1 = vols[r1, j] + wx * (vols[r1, j + 1] - vols[r1, j])
  if outside is not None:
    o = np.asarray(outside, np.float64)
    lo, hi = fraw < 0, fraw > nx - 1
    s0 = np.where(lo, o[row, 0], np.where(hi, o[row, 1], s0))
    s1 = np.where(lo, o[r1, 0], np.where(hi, o[r1, 1], s1))
  return s0 + wt * (s1 - s0)


def lv_fns(times, x0, dx, vols, r, q, outside=None):
  """(drift_fn, volatility_fn) on ln S for oracle.euler."""
  def drift(t, x):
    s = table_sigma(times, x0, dx, vols, float(t), x[..., 0], outside)
    return (r - q - s * s / 2)[..., None]

  def vol(t, x):
    s = table_sigma(times, x0, dx, vols, float(t), x[..., 0], outside)
    return s[..., None, None]
  return drift, vol
