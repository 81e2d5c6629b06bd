"""Oracle: exact-scheme geometric Brownian motion paths.

  * univariate `sample_paths`  tff/models/geometric_brownian_motion/
    univariate_geometric_brownian_motion.py:261-317:
    log increments (mu - sigma^2/2) dt + sqrt(sigma^2 dt) z, cumulative sum
    over the requested times, S = S0 exp(cumsum).
  * multivariate `sample_paths` .../multivariate_geometric_brownian_motion.py:
    229-282: increments (mu - sigma^2/2) dt + sqrt(dt) sigma (L z),
    L = chol(corr).

Draws are injected as the reference's `normal_draws` [N, k, d].
Pinned by: log-moment closed forms E[ln S_t] = ln S0 + (mu - sigma^2/2) t and
Var = sigma^2 t at 3 standard errors (geometric_brownian_motion_test.py:
248-283), the single-step identity S_t = S0 exp((mu - s^2/2) t + s sqrt(t) z)
for one requested time, and correlation of log increments = corr.
"""
import numpy as np


def sample_paths(times, mu, sigma, s0, draws, dtype=np.float64):
  times = np.asarray(times, dtype=dtype)
  z = np.asarray(draws, dtype=dtype)[..., 0]  # [N, k]
  t = np.concatenate([np.zeros(1, dtype), times])
  dt = t[1:] - t[:-1]
  inc = (mu * dt - sigma**2 * dt / 2) + np.sqrt(sigma**2 * dt) * z
  return (s0 * np.exp(np.cumsum(inc, axis=-1)))[..., None].astype(dtype)


def mv_sample_paths(times, means, vols, corr, s0, draws, dtype=np.float64):
  times = np.asarray(times, dtype=dtype)
  z = np.asarray(draws, dtype=dtype)  # [N, k, d]
  t = np.concatenate([np.zeros(1, dtype), times])
  dt = (t[1:] - t[:-1])[None, :, None]
  means = np.asarray(means, dtype)
  vols = np.asarray(vols, dtype)
  if corr is not None:
    chol = np.linalg.cholesky(np.asarray(corr, dtype))
    z = np.einsum('ij,nkj->nki', chol, z)
  inc = (means - vols**2 / 2) * dt + np.sqrt(dt) * vols * z
  return (np.asarray(s0, dtype) * np.exp(np.cumsum(inc, axis=1))).astype(dtype)


def basket_call_paths(s0, sigma, corr, r, t, strike, z, weights=None):
  """Per-path discounted basket-call payoff and its pathwise derivatives.

  z: normals [N, steps, d] (the exact correlated scheme of
  multivariate_geometric_brownian_motion.py:229-282 over uniform steps).
  Returns dict of arrays: pay [N], delta [N, d], vega [N, d], rho [N].
  Pinned by central finite differences of `basket_call_price` with common
  random numbers (tests/test_oracle.py).
  """
  s0 = np.asarray(s0, np.float64)
  sigma = np.asarray(sigma, np.float64)
  d = s0.shape[0]
  w = np.full(d, 1.0 / d) if weights is None else np.asarray(weights)
  n, steps, _ = z.shape
  dt = t / steps
  chol = np.linalg.cholesky(np.asarray(corr, np.float64))
  wv = np.sqrt(dt) * np.einsum('ij,nj->ni', chol, z.sum(axis=1))
  s = s0 * np.exp((r - 0.5 * sigma**2) * t + sigma * wv)
  a = s @ w
  disc = np.exp(-r * t)
  itm = (a > strike).astype(np.float64)
  pay = disc * itm * (a - strike)
  return {'pay': pay,
          'delta': disc * itm[:, None] * w * s / s0,
          'vega': disc * itm[:, None] * w * s * (wv - sigma * t),
          'rho': -t * pay + disc * itm * t * a}


def basket_call_price(s0, sigma, corr, r, t, strike, z, weights=None):
  return basket_call_paths(s0, sigma, corr, r, t, strike, z,
                           weights)['pay'].mean()


def gbm_log_euler_call(s0, sigma, r, t, strike, z):
  """Discounted call payoffs of log-Euler GBM over uniform steps; z [N, steps]."""
  steps = z.shape[1]
  dt = t / steps
  x = np.log(s0) + np.sum((r - 0.5 * sigma**2) * dt + sigma * np.sqrt(dt) * z,
                          axis=1)
  return np.exp(-r * t) * np.maximum(np.exp(x) - strike, 0)
