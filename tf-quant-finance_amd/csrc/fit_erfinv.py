"""Derives the float32 erfinv polynomials used by the fused kernels.

erfinv(x) = x * P(w) with w = -ln(1 - x^2).  Two ranges:
  central: w < 5,  P_c(w - 2.5)           (|x| < 0.99663)
  tail:    w >= 5, P_t(sqrt(w) - 3)       (up to w = 17, i.e. 1 - x^2 >= 4e-8)
Each P is a degree-8 polynomial fitted by weighted least squares on
Chebyshev nodes against scipy.special.erfinv (relative error), then checked
by evaluating the Horner form in float32.  The form of the approximation
(w, the split at w = 5, the variables w - 2.5 and sqrt(w) - 3) follows
M. Giles, "Approximating the erfinv function", GPU Computing Gems Jade
Edition (2011); the coefficients are this script's own fit.
Run: python fit_erfinv.py
"""
import numpy as np
from scipy import special


def fit(lo, hi, tmap, xmap, deg=8, n=4000):
  k = np.arange(n)
  t = 0.5 * (lo + hi) + 0.5 * (hi - lo) * np.cos(np.pi * (k + 0.5) / n)
  x = xmap(t)
  target = special.erfinv(x) / x
  a = np.vander(tmap(t), deg + 1)
  w = 1.0 / np.abs(target)
  c, *_ = np.linalg.lstsq(a * w[:, None], target * w, rcond=None)
  return c  # highest power first


def x_from_w(w):
  return np.sqrt(1.0 - np.exp(-w))


def main():
  cc = fit(1e-12, 5.0, lambda w: w - 2.5, x_from_w)
  ct = fit(np.sqrt(5.0), np.sqrt(17.0), lambda s: s - 3.0,
           lambda s: x_from_w(s * s))
  for name, c in (('central', cc), ('tail', ct)):
    print(name, ', '.join('%.9ef' % v for v in c))
  # float32 check over the whole float32 input range used by the kernels
  q = np.linspace(-1 + 2**-23, 1 - 2**-23, 2_000_001).astype(np.float32)
  # 1 - q*q rounded once, as the kernels' fma(-q, q, 1) does
  w = (-np.log((1 - q.astype(np.float64)**2).astype(np.float32))).astype(
      np.float32)
  p = np.where(w < 5, np.polyval(cc.astype(np.float32), (w - 2.5).astype(np.float32)),
               np.polyval(ct.astype(np.float32), (np.sqrt(w) - 3).astype(np.float32)))
  got = (p * q).astype(np.float64)
  ref = special.erfinv(q.astype(np.float64))
  nz = ref != 0
  print('max rel err', np.max(np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])))


if __name__ == '__main__':
  main()
