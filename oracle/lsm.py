"""Oracle: Longstaff-Schwartz least-squares Monte Carlo.

  * `make_polynomial_basis`  tff/models/longstaff_schwartz/lsm.py:50-125:
    centre by the mean over paths at the date, tensor-product monomials
    with exponents from meshgrid(range(deg+1) x dim) ('xy' indexing), output
    [batch, (deg+1)^dim, N].
  * `least_square_mc`        lsm.py:128-295 (discount padding/transposition,
    terminal payoff as initial cashflow, calibration rows, final mean)
  * `_expected_exercise_fn`  lsm.py:328-383 (ITM mask exercise > 0, masked
    design, X^T X, pinv, X^T y, relu(X beta))
  * `_updated_cashflows_and_values` lsm.py:386-400
  * `_lsm_loop_body`         lsm.py:403-436
  * `make_basket_put_payoff` payoff_utils.py:27-97: relu(K - mean_d S).

numpy.linalg.pinv stands in for tf.linalg.pinv (its rcond default differs:
TF uses 10 max(M, N) eps).  Pinned by the Longstaff-Schwartz (2001) 8-path
table values of lsm_test.py:26-220 (0.0564, 0.1144, 0.174226,
[0.1144, 0.199], [0.1144, 0.1157]) and the duplicated-asset invariant
(lsm_test.py:130-159).
"""
import numpy as np


def exponent_grid(degree, dim):
  """Rows of meshgrid(*(dim*[range(deg+1)])) ('xy') stacked on the last axis."""
  g = np.arange(degree + 1)
  mesh = np.meshgrid(*([g] * dim))  # numpy default indexing='xy' as TF
  return np.stack(mesh, -1).reshape(-1, dim)


def make_polynomial_basis(degree, standardize=False):
  """standardize=True divides the centred values by their standard
  deviation: every column is rescaled by a constant, so the least-squares
  fit is unchanged in exact arithmetic while the Gram matrix becomes well
  conditioned (the form the GPU kernels solve)."""
  def basis(sample_paths, time_index):
    x = np.asarray(sample_paths)
    if x.ndim == 3:
      x = x[None]
    sl = x[:, :, time_index, :]  # [B, N, d]
    c = sl - sl.mean(axis=1, keepdims=True)
    if standardize:
      sd = sl.std(axis=1, keepdims=True)
      c = c / np.where(sd > 0, sd, 1.0)
    e = exponent_grid(degree, x.shape[-1])  # [K, d]
    # prod_j c_j^e_kj, the factors multiplied in asset order as
    # reduce_prod does, from a power table (no [B, N, K, d] intermediate)
    pw = c[..., None]**np.arange(degree + 1).astype(x.dtype)  # [B, N, d, D]
    out = pw[:, :, 0, e[:, 0]]
    for j in range(1, x.shape[-1]):
      out = out * pw[:, :, j, e[:, j]]  # [B, N, K]
    return np.transpose(out, (0, 2, 1))
  return basis


def make_basket_put_payoff(strikes, dtype=None):
  strikes = np.asarray(strikes, dtype=dtype)

  def payoff(sample_paths, time_index):
    x = np.asarray(sample_paths, dtype=strikes.dtype)
    if x.ndim == 3:
      x = x[:, None]
    else:
      x = np.transpose(x, (1, 0, 2, 3))
    avg = x[:, :, time_index, :].mean(-1)  # [N, B]
    return np.maximum(strikes - avg, 0)
  return payoff


def make_basket_call_max_payoff(strikes, dtype=None):
  """relu(max_d S - K): the Bermudan max-call payoff (user payoff_fn)."""
  strikes = np.asarray(strikes, dtype=dtype)

  def payoff(sample_paths, time_index):
    x = np.asarray(sample_paths, dtype=strikes.dtype)
    if x.ndim == 3:
      x = x[:, None]
    else:
      x = np.transpose(x, (1, 0, 2, 3))
    mx = x[:, :, time_index, :].max(-1)
    return np.maximum(mx - strikes, 0)
  return payoff


def _pinv(a):
  n = a.shape[-1]
  rcond = 10 * n * np.finfo(a.dtype).eps
  return np.linalg.pinv(a, rcond=rcond)


def _expected_exercise(design, calib, cont, ex):
  mask = ex > 0  # [N, B]
  dt = np.transpose(design, (0, 2, 1))  # [B, N, K]
  masked = np.where(mask.T[..., None], dt, 0)
  if calib is None:
    sub, y = masked, cont
  else:
    sub, y = masked[:, calib], cont[calib]
  subt = np.transpose(sub, (0, 2, 1))
  lhs = np.matmul(subt, sub)  # X^T X, [B, K, K]
  rhs = np.matmul(subt, np.transpose(y)[..., None])[..., 0]  # X^T y
  beta = np.matmul(_pinv(lhs), rhs[..., None])[..., 0]
  c = np.matmul(dt, beta[..., None])[..., 0]  # [B, N]
  return np.maximum(c.T, 0)


def least_square_mc(sample_paths, exercise_times, payoff_fn, basis_fn,
                    discount_factors=None, num_calibration_samples=None,
                    dtype=None, return_path_values=False):
  """Price per payoff batch entry; with return_path_values also the
  discounted per-path values [N_price, B] whose mean is the price."""
  x = np.asarray(sample_paths, dtype=dtype)
  dtype = x.dtype
  ex_t = list(np.asarray(exercise_times).reshape(-1))
  n_times = len(ex_t)
  if discount_factors is None:
    df = np.ones((1, 1, n_times), dtype)
  else:
    df = np.asarray(discount_factors, dtype)
    if df.ndim == 0:
      df = df.reshape(1, 1, 1)
    elif df.ndim == 1:
      df = df.reshape(1, 1, -1)
    elif df.ndim == 2:
      df = df.reshape(1, -1)
  df = np.pad(df, [(0, 0)] * (df.ndim - 1) + [(1, 0)], constant_values=1)
  df = np.transpose(df, (df.ndim - 1,) + tuple(range(df.ndim - 1)))

  def disc(v, i):
    return (df[i + 1] / df[i]) * v

  cash = payoff_fn(x, ex_t[-1])
  vals = np.zeros_like(cash)
  calib = None
  if num_calibration_samples is not None:
    calib = np.arange(num_calibration_samples)
  e = n_times - 1
  while e > 0:
    ti = ex_t[e - 1]
    exv = payoff_fn(x, ti)
    cont = disc(vals + cash, e)
    design = basis_fn(x, ti)
    exp_cont = _expected_exercise(design, calib, cont, exv)
    upd = exv > exp_cont
    new_vals = np.where(upd, 0, cash + vals)
    cash = np.where(upd, exv, 0)
    vals = disc(new_vals, e)
    e -= 1
  pv = disc(cash + vals, 0)
  if num_calibration_samples is not None:
    pv = pv[num_calibration_samples:]
  if return_path_values:
    return pv.mean(axis=0), pv
  return pv.mean(axis=0)
