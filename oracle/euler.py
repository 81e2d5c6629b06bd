"""Oracle: Euler-Maruyama sampling with injected normal draws.

  * `sample`        tff/models/euler_sampling.py:27-332 (argument handling,
                    grid, draw transposition)
  * `_while_loop`   tff/models/euler_sampling.py:405-464 (record semantics:
                    write at `written_count`, advance on keep_mask[i+1],
                    stop once every requested time is written)
  * `_euler_step`   tff/models/euler_sampling.py:513-537:
                    t = times[i+1]; dw = z_i * sqrt_dt[i];
                    x' = (x + dt[i] * a(t, x)) + b(t, x) @ dw
  * `draw_layout`   tff/models/utils.py:20-128: quasi draws for path p and
                    step i, component k sit at sequence dimension i*d + k.

The drift/volatility callables take numpy arrays [N, d] and return [N, d] /
[N, d, d].  Pinned by closed forms (tests/test_oracle.py): with d=1,
a = mu x, b = sigma x, the Euler path is S0 * prod(1 + mu dt_i + sigma
sqrt(dt_i) z_i) (a product form, not this update); Wiener paths (a = 0,
b = 1) are cumsum(sqrt(dt) z); with b = 0 and a = mu sqrt(t) on a
non-uniform grid the value is exactly the right Riemann sum
x0 + sum dt_i mu sqrt(t_{i+1}) (coefficients at times[i+1]; the left sum
differs by 1e-2); E[X_t] = x0 + 2/3 mu t^1.5 for a = mu sqrt(t),
b = a t + b0 (euler_sampling_test.py:206-294) on the reference's grid.
"""
import numpy as np

from oracle import grid as grid_lib


def step_coefficients(times, dtype):
  times = np.asarray(times, dtype=dtype)
  dt = (times[1:] - times[:-1]).astype(dtype)
  return dt, np.sqrt(dt).astype(dtype)


def run(drift_fn, volatility_fn, times_grid, keep_mask, num_requested,
        x0, draws, dtype):
  """Runs `_while_loop` with draws [steps, N, d]; returns [N, k, d]."""
  all_times = np.asarray(times_grid, dtype=dtype)
  dt, sqrt_dt = step_coefficients(all_times, dtype)
  steps = dt.shape[0]
  x = np.array(x0, dtype=dtype, copy=True)
  n, d = x.shape
  record = num_requested != 1
  out = np.zeros((num_requested, n, d), dtype=dtype)
  written = 0
  if record:
    out[0] = x
  written += int(keep_mask[0])
  i = 0
  while i < steps and written < num_requested:
    t = all_times[i + 1]
    dw = (draws[i] * sqrt_dt[i]).astype(dtype)
    inc = (dt[i] * drift_fn(t, x)).astype(dtype)
    vol = volatility_fn(t, x).astype(dtype)
    dw_inc = np.einsum('nij,nj->ni', vol, dw).astype(dtype)
    x = ((x + inc) + dw_inc).astype(dtype)
    if record:
      out[written] = x
    else:
      out[0] = x
    written += int(keep_mask[i + 1])
    i += 1
  return np.transpose(out, (1, 0, 2))


def sample(dim, drift_fn, volatility_fn, times, num_samples=1,
           initial_state=None, normal_draws=None, num_time_steps=None,
           time_step=None, times_grid=None, dtype=np.float64,
           tolerance=None):
  """Mirror of euler_sampling.sample with `normal_draws` [N, steps, d]."""
  times = np.asarray(times, dtype=dtype)
  all_times, mask, _ = grid_lib.prepare_grid(
      times, dtype, time_step=time_step, num_time_steps=num_time_steps,
      times_grid=times_grid, tolerance=tolerance)
  draws = np.transpose(np.asarray(normal_draws, dtype=dtype), (1, 0, 2))
  n = draws.shape[1]
  if initial_state is None:
    initial_state = np.zeros(dim, dtype)
  x0 = np.asarray(initial_state, dtype=dtype) + np.zeros((n, dim), dtype)
  return run(drift_fn, volatility_fn, all_times, mask, times.shape[0], x0,
             draws, dtype)


def quasi_draw_layout(flat, num_steps, dim):
  """[N, steps*d] sequence points -> [steps, N, d] (models/utils.py:114-127)."""
  n = flat.shape[0]
  return np.transpose(flat.reshape(n, num_steps, dim), (1, 0, 2))


def heston_ft_fns(kappa, theta, xi, rho, r):
  """Full-truncation Heston drift/vol in the reference's callable format.

  Coefficients follow heston_model.py:143-173 with v -> v+ = max(v, 0) and
  the log-spot drift r - v+/2.
  """
  def drift(t, x):
    vp = np.maximum(x[:, 1], 0)
    return np.stack([r - 0.5 * vp, kappa * (theta - vp)], -1)

  def vol(t, x):
    vp = np.maximum(x[:, 1], 0)
    s = np.sqrt(vp)
    z = np.zeros_like(s)
    c1 = np.stack([s, xi * rho * s], -1)
    c2 = np.stack([z, xi * np.sqrt(1 - rho**2) * s], -1)
    return np.stack([c1, c2], -1)
  return drift, vol


def gbm_fns(mu, sigma):
  """GeometricBrownianMotion drift/vol callables (univariate...py:127-153)."""
  def drift(t, x):
    return mu * x

  def vol(t, x):
    return sigma * x[..., None]
  return drift, vol


def gbm_log_fns(mu, sigma):
  """GBM in log space, the notebook's GenericItoProcess setup."""
  def drift(t, x):
    return np.full_like(x, mu - 0.5 * sigma**2)

  def vol(t, x):
    return np.full(x.shape + (1,), sigma, dtype=x.dtype)
  return drift, vol
