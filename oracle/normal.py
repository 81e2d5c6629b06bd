"""Oracle: uniform -> normal transforms.

  * quasi draws: z = erfinv((u - 0.5) * 2) * sqrt(2)
    (tff/math/random_ops/multivariate_normal.py:420), evaluated here in
    float64 with scipy.special.erfinv.
  * the engine's shifted-Sobol transform: float32 uses the 25-bit uniform
    u = (n + 1/2) 2^-25 of the top 25 bits n of a 32-bit point, float64 the
    32-bit u = (v + 1/2) 2^-32; q = 2u - 1 is formed exactly and
    z = sqrt(2) erfinv(q).

Pinned by scipy's independent special functions (erfinv/ndtri agree) and by
symmetry z(u) = -z(1-u); the device float32 transform is compared with a
stated ULP-style tolerance, the float64 one at 1e-13 relative.
"""
import numpy as np
from scipy import special


def quasi_normals(u):
  u = np.asarray(u, np.float64)
  return special.erfinv((u - 0.5) * 2) * np.sqrt(2.0)


def shifted_sobol_normals(ints_u32, dtype=np.float32):
  """float32 kernels use the top 25 bits, float64 kernels all 32 bits."""
  v = np.asarray(ints_u32, np.uint64)
  if dtype == np.float64:
    q = (v.astype(np.float64) + 0.5) * 2.0**-31 - 1.0
  else:
    n = (v >> np.uint64(7)).astype(np.float64)
    q = (2 * n + 1 - 2.0**25) * 2.0**-25
  return special.erfinv(q) * np.sqrt(2.0)
