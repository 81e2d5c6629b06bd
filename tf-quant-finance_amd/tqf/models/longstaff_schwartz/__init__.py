"""Longstaff-Schwartz least-squares Monte Carlo (mirror of
tff.models.longstaff_schwartz).

  least_square_mc           lsm.py:128-295 (+ loop body 298-436)
  make_polynomial_basis     lsm.py:50-125
  make_basket_put_payoff    payoff_utils.py:27-97
  make_max_call_payoff      engine extension: relu(max_j S_j - K)
  make_table_payoff         engine extension: exercise values given per
                            (path, payoff, date), e.g. swap values

With the basis and payoff built here, the whole backward induction runs in
the tqf_lsm_price kernels (float64 normal equations built on the fly,
Cholesky solve on the GPU).  The kernels standardise each asset before
forming the monomials and Jacobi-scale the normal equations, so they solve
the least-squares problem exactly where the reference's pinv of the raw
tensor-product Gram (condition numbers past 1e15 for degree >= 2 on three
or more assets) discards directions; on well-conditioned problems the two
agree.  Other callables run the reference's algorithm as torch ops on the
same GPU (torch.linalg.pinv with TF's cutoff).
"""
import ctypes

import numpy as np
import torch

from tqf import _native
from tqf.models import utils

__all__ = ['least_square_mc', 'make_polynomial_basis',
           'make_basket_put_payoff', 'make_max_call_payoff',
           'make_table_payoff', 'exponent_grid']

PAYOFF_BASKET_PUT, PAYOFF_MAX_CALL, PAYOFF_TABLE = 0, 1, 2


def exponent_grid(degree, dim):
  """Exponents [K, dim] in the order of tf.meshgrid(*dim*[range(deg+1)])
  ('xy' indexing) stacked on the last axis (lsm.py:116-118)."""
  g = np.arange(degree + 1)
  mesh = np.meshgrid(*([g] * dim), indexing='xy')
  return np.stack(mesh, -1).reshape(-1, dim)


def make_polynomial_basis(degree):
  """Centred tensor-product monomials, output [batch, (deg+1)^dim, N]."""
  degree = int(degree)

  def basis(sample_paths, time_index):
    x = torch.as_tensor(sample_paths)
    if x.dim() == 3:
      x = x.unsqueeze(0)
    sl = x[:, :, int(time_index), :]
    c = sl - sl.mean(dim=1, keepdim=True)
    e = torch.as_tensor(exponent_grid(degree, x.shape[-1]), dtype=x.dtype,
                        device=x.device)
    out = torch.prod(c.unsqueeze(2)**e, dim=-1)  # [B, N, K]
    return out.permute(0, 2, 1)
  basis._tqf_basis = ('poly', degree)
  return basis


def _payoff(kind, strikes, dtype):
  strikes_np = np.asarray(utils.to_numpy(strikes), np.float64).reshape(-1)

  def payoff(sample_paths, time_index):
    x = torch.as_tensor(sample_paths)
    if x.dim() == 3:
      x = x.unsqueeze(1)
    else:
      x = x.permute(1, 0, 2, 3)
    sl = x[:, :, int(time_index), :]
    k = torch.as_tensor(strikes_np, dtype=x.dtype, device=x.device)
    if kind == PAYOFF_BASKET_PUT:
      return torch.relu(k - sl.mean(-1))
    return torch.relu(sl.max(-1).values - k)
  payoff._tqf_payoff = (kind, strikes_np)
  return payoff


def make_basket_put_payoff(strikes, dtype=None, name=None):
  """relu(K - mean_j S_j) per payoff batch entry."""
  del name
  return _payoff(PAYOFF_BASKET_PUT, strikes, dtype)


def make_max_call_payoff(strikes, dtype=None, name=None):
  """relu(max_j S_j - K): the Bermudan max-call payoff."""
  del name
  return _payoff(PAYOFF_MAX_CALL, strikes, dtype)


def make_table_payoff(values):
  """payoff_fn(paths, t) = values[:, :, t] for exercise values [N, B,
  n_times] computed elsewhere (non-negative; in the money iff > 0)."""
  table = torch.as_tensor(values)

  def payoff(sample_paths, time_index):
    del sample_paths
    return table[:, :, int(time_index)]
  payoff._tqf_payoff = (PAYOFF_TABLE, table)
  return payoff


def _padded_discounts(discount_factors, n_ex, dtype):
  """Reference padding/transposition (lsm.py:239-256) -> [n_ex + 1, *, B].
  A torch tensor stays a tensor on its device (per-path factors are as
  large as the paths); anything else becomes a numpy array."""
  if torch.is_tensor(discount_factors):
    df = discount_factors.to(torch.float64)
    if df.ndim == 0:
      df = df.reshape(1, 1, 1)
    elif df.ndim == 1:
      df = df.reshape(1, 1, -1)
    elif df.ndim == 2:
      df = df.reshape(1, -1)
    df = torch.nn.functional.pad(df, (1, 0), value=1.0)
    return df.permute((df.ndim - 1,) + tuple(range(df.ndim - 1)))
  if discount_factors is None:
    df = np.ones((1, 1, n_ex))
  else:
    df = np.asarray(utils.to_numpy(discount_factors), np.float64)
    if df.ndim == 0:
      df = df.reshape(1, 1, 1)
    elif df.ndim == 1:
      df = df.reshape(1, 1, -1)
    elif df.ndim == 2:
      df = df.reshape(1, -1)
  df = np.pad(df, [(0, 0)] * (df.ndim - 1) + [(1, 0)], constant_values=1)
  return np.transpose(df, (df.ndim - 1,) + tuple(range(df.ndim - 1)))


def least_square_mc(sample_paths, exercise_times, payoff_fn, basis_fn,
                    discount_factors=None, num_calibration_samples=None,
                    dtype=None, name=None):
  """Price of an American/Bermudan option per payoff batch entry."""
  del name
  dev = _native.device()
  if dtype is not None and not isinstance(dtype, torch.dtype):
    dtype = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
  x = torch.as_tensor(sample_paths, device=dev)
  if dtype is not None:
    x = x.to(dtype)
  elif not x.is_floating_point():
    x = x.to(torch.float32)
  dtype = x.dtype
  ex = [int(v) for v in np.asarray(utils.to_numpy(exercise_times)).reshape(-1)]
  n_ex = len(ex)
  df = _padded_discounts(discount_factors, n_ex, dtype)
  bspec = getattr(basis_fn, '_tqf_basis', None)
  pspec = getattr(payoff_fn, '_tqf_payoff', None)
  # discount factors are one per date, or (rank 3) one per path and date;
  # other broadcast shapes run through the torch restatement
  n_paths = x.shape[-3]
  df_ok = df.ndim <= 2 or df.shape[1] in (1, n_paths)
  if (bspec is not None and pspec is not None and bspec[0] == 'poly' and
      df_ok):
    return _native_lsm(x, ex, pspec, bspec[1], df, num_calibration_samples,
                       dtype, dev)
  return _torch_lsm(x, ex, payoff_fn, basis_fn, df, num_calibration_samples,
                    dtype)


def step_major_f64(paths):
  """[N, T, d] paths as a float64 view whose memory is [T][d][N] (the
  kernels read one date of every path per pass: with this layout those
  reads are coalesced).  Returns the input when it already is one."""
  if paths.dtype == torch.float64 and paths.stride(0) == 1 and (
      paths.stride(2) == paths.shape[0] and
      paths.stride(1) == paths.shape[0] * paths.shape[2]):
    return paths
  n, t, d = paths.shape
  out = torch.empty((t, d, n), dtype=torch.float64, device=paths.device)
  out.copy_(paths.permute(1, 2, 0))
  return out.permute(2, 0, 1)


def _native_lsm(x, ex, pspec, degree, df, num_calib, dtype, dev,
                return_details=False, basis_log=False, exponents=None,
                path_values=None):
  """basis_log: False / 0 regresses on the path values, True / 1 on their
  logarithms, 2 takes paths that hold log prices (basis on them, payoff on
  their exp).
  exponents: an optional [K, d] table of monomial exponents (each at
  most `degree`) in place of the tensor grid of make_polynomial_basis.
  path_values: an optional float64 device tensor [N] that receives, for the
  last payoff batch entry, every path's value at the first exercise date
  (cashflow + discounted later values, before the final discount df[1] /
  df[0]): the per-path terms whose mean over the pricing rows is the
  price."""
  kind, strikes = pspec
  if int(basis_log) not in (0, 1, 2):
    raise ValueError('basis_log must be 0, 1 or 2')
  paths_b = x.unsqueeze(0) if x.dim() == 3 else x  # [Bp, N, T, d]
  bp = paths_b.shape[0]
  table = None
  if kind == PAYOFF_TABLE:
    table = strikes.to(device=dev, dtype=torch.float64)  # [N, B, T]
    strikes = np.ones(table.shape[1])
  n_batch = max(bp, strikes.shape[0], df.shape[-1])
  per_path = df.ndim == 3 and df.shape[1] > 1
  d = paths_b.shape[-1]
  if exponents is None:
    expo = exponent_grid(degree, d).astype(np.uint8)
  else:
    expo = np.ascontiguousarray(exponents, np.uint8)
    if expo.ndim != 2 or expo.shape[1] != d or expo.max() > degree:
      raise ValueError('exponent table must be [K, %d] with entries <= %d'
                       % (d, degree))
  k = expo.shape[0]
  if k > 255 or d > 8:
    raise NotImplementedError('the GPU LSM supports at most 8 assets and 255 '
                              'basis functions')
  expo_t = torch.as_tensor(expo, device=dev).contiguous()
  n = paths_b.shape[1]
  ws = torch.empty(int(_native.lib().tqf_lsm_workspace_size(n, k)) // 8 + 1,
                   dtype=torch.float64, device=dev)
  cash = torch.empty(n, dtype=torch.float64, device=dev)
  vals = torch.empty(n, dtype=torch.float64, device=dev)
  exi = (ctypes.c_int32 * len(ex))(*ex)
  prices, ses, ranks_all = [], [], []
  for b in range(n_batch):
    p64 = step_major_f64(paths_b[b if bp > 1 else 0])
    l = _native.Lsm()
    l.paths = p64.data_ptr()
    l.stride_path, l.stride_time, l.stride_asset = p64.stride()
    l.n_paths = n
    l.n_times = p64.shape[1]
    l.dim = d
    l.payoff = kind
    l.degree = degree
    l.n_basis = k
    l.basis_log = int(basis_log)
    l.exponents = expo_t.data_ptr()
    l.strike = float(strikes[b if strikes.shape[0] > 1 else 0])
    l.num_calibration = int(num_calib or 0)
    l.rcond = 10 * k * np.finfo(np.float64).eps
    keep = []
    if table is not None:
      tb = table[:, b if table.shape[1] > 1 else 0, :].T.contiguous()
      keep.append(tb)
      l.payoff_table = tb.data_ptr()
    if per_path:
      col = df[:, :, b if df.shape[2] > 1 else 0]  # [n_ex + 1, N]
      if torch.is_tensor(col):
        pdb = col.to(device=dev, dtype=torch.float64).contiguous()
      else:
        pdb = torch.as_tensor(np.ascontiguousarray(col), dtype=torch.float64,
                              device=dev).contiguous()
      keep.append(pdb)
      l.path_discounts = pdb.data_ptr()
      dfb = np.ones(df.shape[0])
    else:
      dfn = df.cpu().numpy() if torch.is_tensor(df) else df
      dfb = np.ascontiguousarray(
          np.broadcast_to(dfn, dfn.shape[:-1] + (n_batch,))[..., b].reshape(
              dfn.shape[0], -1)[:, 0], np.float64)
    res = torch.empty(2, dtype=torch.float64, device=dev)
    ranks = torch.zeros(max(len(ex) - 1, 1), dtype=torch.int32, device=dev)
    _native.call('tqf_lsm_price', ctypes.byref(l), exi,
                 ctypes.c_int32(len(ex)),
                 dfb.ctypes.data_as(ctypes.c_void_p),
                 ctypes.c_void_p(ws.data_ptr()),
                 ctypes.c_void_p(cash.data_ptr()),
                 ctypes.c_void_p(vals.data_ptr()),
                 ctypes.c_void_p(res.data_ptr()),
                 ctypes.c_void_p(ranks.data_ptr()), _native.stream_ptr(dev))
    prices.append(res[0])
    ses.append(res[1])
    ranks_all.append(ranks)
    if path_values is not None:
      torch.add(cash, vals, out=path_values)
  out = torch.stack(prices).to(dtype)
  if return_details:
    return out, torch.stack(ses), torch.stack(ranks_all)
  return out


def _torch_lsm(x, ex, payoff_fn, basis_fn, df, num_calib, dtype):
  """The reference algorithm (lsm.py:258-436) as torch ops on the GPU."""
  dev = x.device
  dft = torch.as_tensor(df, dtype=dtype, device=dev)

  def disc(v, i):
    return (dft[i + 1] / dft[i]) * v

  n_ex = len(ex)
  cash = payoff_fn(x, ex[-1])
  vals = torch.zeros_like(cash)
  calib = None
  if num_calib is not None:
    calib = torch.arange(num_calib, device=dev)
  e = n_ex - 1
  while e > 0:
    ti = ex[e - 1]
    exv = payoff_fn(x, ti)
    cont = disc(vals + cash, e)
    design = basis_fn(x, ti)  # [B, K, N]
    mask = exv > 0
    dt = design.permute(0, 2, 1)
    masked = torch.where(mask.T.unsqueeze(-1), dt, torch.zeros_like(dt))
    if calib is None:
      sub, y = masked, cont
    else:
      sub, y = masked[:, calib], cont[calib]
    lhs = torch.matmul(sub.transpose(1, 2).contiguous(), sub)
    rhs = torch.matmul(sub.transpose(1, 2).contiguous(),
                       y.T.unsqueeze(-1))
    # tf.linalg.pinv cutoff: 10 * max(M, N) * eps of the largest singular
    # value (lsm.py:372)
    rtol = 10 * lhs.shape[-1] * torch.finfo(lhs.dtype).eps
    beta = torch.matmul(torch.linalg.pinv(lhs, rtol=rtol), rhs)
    c = torch.relu(torch.matmul(dt, beta).squeeze(-1).T)
    upd = exv > c
    new_vals = torch.where(upd, torch.zeros_like(cash), cash + vals)
    cash = torch.where(upd, exv, torch.zeros_like(cash))
    vals = disc(new_vals, e)
    e -= 1
  pv = disc(cash + vals, 0)
  if num_calib is not None:
    pv = pv[num_calib:]
  return pv.mean(dim=0)
