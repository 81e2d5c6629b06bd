"""Oracle: Heston model.

  * `drift_vol`            tff/models/heston/heston_model.py:143-173 (Euler
                           coefficients: [-v/2, k(theta - v)] and
                           [[sqrt|v|, 0], [xi rho sqrt|v|, xi sqrt(1-rho^2) sqrt|v|]])
  * `update_variance`      heston_model.py:522-551 (Andersen QE, psi_c = 1.5)
  * `update_log_spot`      heston_model.py:554-572 (gamma_1 = gamma_2 = 1/2)
  * `sample_paths_qe`      heston_model.py:177-460 (grid, draws [steps, N, 2]
                           with channel 0 -> variance, 1 -> log spot)
  * `expected_total_variance` heston_model.py:462-508
  * `european_call`        an independent semi-analytic price: Lewis (2001)
                           single-integral formula with the Albrecher et al.
                           (2007) characteristic function, scipy quadrature.

Pinned by: the reference's coefficient goldens (heston_model_test.py:211-219:
drift [-0.0225, 0.955], vol [[0.21213203, 0], [0.00848528, 0.01944222]]),
the expected-total-variance closed form (heston_model_test.py:79-95),
QE moment matching (the QE step reproduces the exact conditional mean of v to
MC error), xi -> 0 reducing to GBM with deterministic variance, the
semi-analytic price against Black-Scholes when xi -> 0, and the QE log-spot
update with its correlation terms: at rho = -0.7, xi = 0.6 the QE call
prices over the smile and the martingale mean match `european_call` within
3 SE, while a sign error planted in those terms misses by more than 10 SE.
`ft_euler_call_greeks` is pinned by central differences of
`ft_euler_calls` on the same draws.
"""
import numpy as np
from scipy import integrate, special

from oracle import grid as grid_lib

_SQRT_2 = np.sqrt(2.0)


def drift_vol(x, kappa, theta, xi, rho):
  v = x[..., 1]
  drift = np.stack([-v / 2, kappa * (theta - v)], -1)
  s = np.sqrt(np.abs(v))
  z = np.zeros_like(s)
  c1 = np.stack([s, xi * rho * s], -1)
  c2 = np.stack([z, xi * np.sqrt(1 - rho**2) * s], -1)
  return drift, np.stack([c1, c2], -1)


def update_variance(kappa, theta, xi, v, dt, z, psi_c=1.5):
  e = np.exp(-kappa * dt)
  xi2 = xi**2
  m = theta + (v - theta) * e
  s2 = (v * xi2 * e / kappa * (1 - e) + theta * xi2 / 2 / kappa * (1 - e)**2)
  psi = s2 / m**2
  u = 0.5 * (1 + special.erf(z / _SQRT_2))
  cond = psi < psi_c
  with np.errstate(divide='ignore', invalid='ignore'):
    psi_inv = 2 / psi
    b2 = psi_inv - 1 + np.sqrt(psi_inv * (psi_inv - 1))
    a = m / (1 + b2)
    v_true = a * (np.sqrt(b2) + z)**2
    p = (psi - 1) / (psi + 1)
    beta = (1 - p) / m
    v_false = np.where(u > p, np.log(1 - p) - np.log(1 - u), 0.0) / beta
  return np.where(cond, v_true, v_false)


def update_log_spot(kappa, theta, xi, rho, v, v_next, x, dt, z,
                    g1=0.5, g2=0.5):
  k0 = -rho * kappa * theta / xi * dt
  k1 = g1 * dt * (kappa * rho / xi - 0.5) - rho / xi
  k2 = g2 * dt * (kappa * rho / xi - 0.5) + rho / xi
  k3 = g1 * dt * (1 - rho**2)
  k4 = g2 * dt * (1 - rho**2)
  return (x + k0 + k1 * v + k2 * v_next + np.sqrt(k3 * v + k4 * v_next) * z)


def sample_paths_qe(times, x0, v0, kappa, theta, xi, rho, draws,
                    num_time_steps=None, time_step=None, tolerance=1e-6,
                    dtype=np.float64):
  """QE paths with draws [N, steps, 2] -> [N, k, 2] (log S, v)."""
  times = np.asarray(times, dtype=dtype)
  all_times, mask = grid_lib.heston_prepare_grid(
      times, dtype, time_step=time_step, num_time_steps=num_time_steps)
  dt = (all_times[1:] - all_times[:-1]).astype(dtype)
  z = np.transpose(np.asarray(draws, dtype=dtype), (1, 0, 2))
  n = z.shape[1]
  k = times.shape[0]
  x = np.full(n, x0, dtype=dtype)
  v = np.full(n, v0, dtype=dtype)
  out_x = np.zeros((k, n), dtype)
  out_v = np.zeros((k, n), dtype)
  record = k != 1
  written = 0
  if record:
    out_x[0], out_v[0] = x, v
  i = 0
  while i < dt.shape[0] and written < k:
    if dt[i] > tolerance:
      v_next = update_variance(kappa, theta, xi, v, dt[i], z[i, :, 0])
      x_next = update_log_spot(kappa, theta, xi, rho, v, v_next, x, dt[i],
                               z[i, :, 1])
    else:
      v_next, x_next = v, x
    v, x = v_next.astype(dtype), x_next.astype(dtype)
    if record:
      out_x[written], out_v[written] = x, v
    else:
      out_x[0], out_v[0] = x, v
    written += int(mask[i + 1])
    i += 1
  return np.stack([out_x.T, out_v.T], -1)


def expected_total_variance(future_times, v0, kappa, theta):
  t = np.asarray(future_times, dtype=np.float64)
  strength = (1 - np.exp(-kappa * t)) / kappa
  return (v0 - theta) * strength + theta * t


def _char_fn(u, t, v0, kappa, theta, xi, rho):
  """E[exp(i u ln(S_T/F))] in the Albrecher 'little trap' form."""
  d = np.sqrt((rho * xi * 1j * u - kappa)**2 + xi**2 * (1j * u + u**2))
  g = (kappa - rho * xi * 1j * u - d) / (kappa - rho * xi * 1j * u + d)
  e = np.exp(-d * t)
  c = kappa * theta / xi**2 * ((kappa - rho * xi * 1j * u - d) * t
                               - 2 * np.log((1 - g * e) / (1 - g)))
  dd = (kappa - rho * xi * 1j * u - d) / xi**2 * (1 - e) / (1 - g * e)
  return np.exp(c + dd * v0)


def european_call(s0, k, t, r, v0, kappa, theta, xi, rho, q=0.0):
  """Lewis (2001): C = S e^{-qT} - sqrt(S K) e^{-(r+q)T/2}/pi *
  int_0^inf Re[e^{i u k} phi(u - i/2)] / (u^2 + 1/4) du, k = ln(F/K)."""
  f = s0 * np.exp((r - q) * t)
  lk = np.log(f / k)

  def integrand(u):
    phi = _char_fn(u - 0.5j, t, v0, kappa, theta, xi, rho)
    return (np.exp(1j * u * lk) * phi).real / (u * u + 0.25)

  val, _ = integrate.quad(integrand, 0.0, np.inf, limit=400, epsabs=1e-12,
                          epsrel=1e-10)
  return np.exp(-r * t) * (f - np.sqrt(f * k) / np.pi * val)


def shifted_sobol_ints(ndims, num_paths, shift, first=0):
  """32-digit Sobol integers of indices first..first+N-1 XOR shift.

  (qmc digital_net_sample with num_digits = 32, digital_net.py:351-407.)
  Returns [N, ndims] int64.
  """
  from oracle import sobol as osobol
  total = max(first + num_paths, 2)
  gm = osobol.sobol_generating_matrices(ndims, total, 32)
  return osobol.digital_net_sample_int(
      gm, total, 32, sequence_indices=np.arange(first, first + num_paths),
      digital_shift=np.asarray(shift, np.int64))


def ft_euler_calls(x0, v0, kappa, theta, xi, rho, r, strikes, all_times,
                   keep_mask, z, dtype=np.float64):
  """Full-truncation log-Euler Heston (euler_sampling.py:513-537 with the
  coefficients of heston_model.py:143-173, v -> max(v, 0), drift r) and
  discounted call payoffs at the kept times.

  z: normals [N, steps, 2].  Vectorised over paths, looping over steps as
  the reference's while_loop does.  Returns (prices [k, K], stderrs [k, K]).
  """
  t = np.asarray(all_times, dtype)
  dt = (t[1:] - t[:-1]).astype(dtype)
  sdt = np.sqrt(dt).astype(dtype)
  n = z.shape[0]
  x = np.full(n, x0, dtype)
  v = np.full(n, v0, dtype)
  xr = dtype(xi * rho)
  xs = dtype(xi * np.sqrt(1 - rho * rho))
  strikes = np.asarray(strikes, np.float64)
  prices, ses = [], []
  for i in range(dt.shape[0]):
    vp = np.maximum(v, 0)
    sq = np.sqrt(vp)
    dw0 = z[:, i, 0].astype(dtype) * sdt[i]
    dw1 = z[:, i, 1].astype(dtype) * sdt[i]
    xn = (x + dt[i] * (dtype(r) - dtype(0.5) * vp)) + sq * dw0
    v = (v + dt[i] * (dtype(kappa) * (dtype(theta) - vp))) + (
        (xr * sq) * dw0 + (xs * sq) * dw1)
    x = xn
    if keep_mask[i + 1]:
      s = np.exp(x.astype(np.float64))
      pay = np.maximum(s[:, None] - strikes[None, :], 0)
      df = np.exp(-r * float(t[i + 1]))
      prices.append(df * pay.mean(0))
      ses.append(df * pay.std(0, ddof=1) / np.sqrt(n))
  return np.array(prices), np.array(ses)


def ft_euler_call_greeks(x0, v0, kappa, theta, xi, rho, r, strike, all_times,
                         z, dtype=np.float64):
  """Price and pathwise Greeks of one discounted call under the
  full-truncation log-Euler Heston scheme (`ft_euler_calls`), by forward
  tangents carried along the step loop, as custom_loops.for_loop carries
  the Jacobian of the state (tff/math/custom_loops.py:20-214, used by
  euler_sampling.py:467-510 with watch_params).

  z: normals [N, steps, 2].  Returns (means, stderrs), each [8]: price,
  d/dS0, d/dv0, d/dr, d/dkappa, d/dtheta, d/dxi, d/drho (S0 = exp(x0)).
  Pinned by central finite differences of `ft_euler_calls` on the same
  draws (tests/test_oracle.py).
  """
  t = np.asarray(all_times, dtype)
  dt = (t[1:] - t[:-1]).astype(dtype)
  sdt = np.sqrt(dt).astype(dtype)
  n = z.shape[0]
  x = np.full(n, x0, dtype)
  v = np.full(n, v0, dtype)
  s1r = np.sqrt(1 - rho * rho)
  xr, xs = xi * rho, xi * s1r
  # tangent rows: x0, v0, r, kappa, theta, xi, rho
  tx = np.zeros((7, n), dtype)
  tv = np.zeros((7, n), dtype)
  tx[0] = 1.0
  tv[1] = 1.0
  for i in range(dt.shape[0]):
    pos = v > 0
    vp = np.where(pos, v, 0)
    sq = np.sqrt(vp)
    dw0 = z[:, i, 0] * sdt[i]
    dw1 = z[:, i, 1] * sdt[i]
    with np.errstate(divide='ignore'):
      hs = np.where(sq > 0, 0.5 / np.where(sq > 0, sq, 1), 0)
    tvp = np.where(pos, tv, 0)
    tsq = tvp * hs
    ntx = tx - dt[i] * 0.5 * tvp + tsq * dw0
    ntv = tv - dt[i] * kappa * tvp + tsq * (xr * dw0 + xs * dw1)
    ntx[2] += dt[i]                                     # r
    ntv[3] += dt[i] * (theta - vp)                      # kappa
    ntv[4] += dt[i] * kappa                             # theta
    ntv[5] += sq * (rho * dw0 + s1r * dw1)              # xi
    ntv[6] += xi * sq * (dw0 - rho / s1r * dw1)         # rho
    x = (x + dt[i] * (r - 0.5 * vp)) + sq * dw0
    v = (v + dt[i] * (kappa * (theta - vp))) + (xr * sq * dw0 + xs * sq * dw1)
    tx, tv = ntx, ntv
  s = np.exp(x.astype(np.float64))
  T = float(t[-1])
  disc = np.exp(-r * T)
  itm = s > strike
  pay = disc * np.where(itm, s - strike, 0)
  xb = disc * np.where(itm, s, 0)
  g = np.empty((8, n))
  g[0] = pay
  g[1] = xb * tx[0] / np.exp(x0)
  g[2] = xb * tx[1]
  g[3] = xb * tx[2] - T * pay
  g[4:] = xb * tx[3:]
  return g.mean(1), g.std(1, ddof=1) / np.sqrt(n)
