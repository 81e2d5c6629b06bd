"""Bermudan swaptions under one-factor Hull-White, priced by
Longstaff-Schwartz on simulated short rates (mirror of
tff.models.hull_white.bermudan_swaption_price,
tff/models/hull_white/swaption.py:314-724; payoff mapping :876-935).

The short rate runs in the fused kernel (HullWhiteModel1F); at every
exercise date the swap's value per path comes from zero-coupon bonds
reconstituted from r(t) (fixed leg sum_j c_j tau_j P(t, T_j), floating leg
1 - P(t, T_n)); the path discount factor is prod exp(-r_k dt_k) on the
simulation grid (the rate at the end of each interval, as the reference
does); the regression is on a polynomial basis of r(t) at the union of all
swaptions' exercise dates, every swaption of the batch one payoff column.
"""
import numpy as np
import torch

from tqf.models import longstaff_schwartz as lsm
from tqf.models import utils
from tqf.models.hull_white.one_factor import HullWhiteModel1F

__all__ = ['bermudan_swaption_price']


def _unique_in_order(x):
  """tf.unique: distinct values in first-appearance order and the index of
  every element into them."""
  vals, first, inv = np.unique(x, return_index=True, return_inverse=True)
  order = np.argsort(first, kind='stable')
  rank = np.empty_like(order)
  rank[order] = np.arange(order.shape[0])
  return vals[order], rank[inv]


def bermudan_swaption_price(*, exercise_times, floating_leg_start_times,
                            floating_leg_end_times, fixed_leg_payment_times,
                            floating_leg_daycount_fractions,
                            fixed_leg_daycount_fractions, fixed_leg_coupon,
                            reference_rate_fn, mean_reversion, volatility,
                            notional=None, is_payer_swaption=True,
                            use_finite_difference=False, lsm_basis=None,
                            num_samples=100, random_type=None, seed=None,
                            skip=0, time_step=None,
                            time_step_finite_difference=None,
                            num_grid_points_finite_difference=100,
                            dtype=None, name=None):
  """Prices [batch] (or a scalar without batch) of Bermudan swaptions with
  exercise_times batch + [n_ex] and fixed_leg_payment_times batch +
  [n_ex, n_pay] (the underlying swap entered at each exercise date)."""
  del floating_leg_start_times, floating_leg_end_times
  del floating_leg_daycount_fractions, time_step_finite_difference
  del num_grid_points_finite_difference, name
  if use_finite_difference:
    raise NotImplementedError(
        'the finite-difference (PDE) valuation is not part of this build; '
        'use the LSM valuation')
  if dtype is None:
    tdt = torch.float64
  elif isinstance(dtype, torch.dtype):
    tdt = dtype
  else:
    tdt = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
  ex = np.asarray(utils.to_numpy(exercise_times), np.float64)
  pay = np.asarray(utils.to_numpy(fixed_leg_payment_times), np.float64)
  dcf = np.asarray(utils.to_numpy(fixed_leg_daycount_fractions), np.float64)
  cpn = np.asarray(utils.to_numpy(fixed_leg_coupon), np.float64)
  notional = 1.0 if notional is None else np.asarray(
      utils.to_numpy(notional), np.float64)
  # as in the reference (swaption.py:587-589), every swaption is priced as
  # a payer swaption
  del is_payer_swaption
  if ex.ndim < pay.ndim - 1:
    raise ValueError('Swaption exercise times not specified for all '
                     'swaptions in the batch. Expected rank '
                     '{} but received {}.'.format(pay.ndim - 1, ex.ndim))
  if time_step is None:
    raise ValueError('`time_step` must be provided for LSM valuation.')
  plan = _plan(ex, pay, time_step)
  model = HullWhiteModel1F(mean_reversion, volatility,
                           initial_discount_rate_fn=reference_rate_fn,
                           dtype=tdt)
  r_t = model.sample_paths(plan['sim_times'], num_samples=num_samples,
                           random_type=random_type, seed=seed, skip=skip)
  bonds = _exercise_bonds(model, plan, r_t, pay)
  basis = lsm.make_polynomial_basis(2) if lsm_basis is None else lsm_basis
  value = _price_on_paths(plan, bonds, r_t, pay, dcf, cpn, basis, tdt)
  value = torch.as_tensor(notional, dtype=tdt, device=r_t.device) * value
  if plan['batch_shape']:
    return value.reshape(plan['batch_shape'])
  return value.reshape(())


def _plan(ex, pay, time_step):
  """Simulation grid and index maps (swaption.py:591-650): the union of
  exercise dates (first-appearance order) and a time_step grid, and the
  sorted distinct bond tenors tau = T_j - t_e."""
  batch_shape = list(ex.shape[:-1])
  unique_ex, ex_index = _unique_in_order(ex.reshape(-1))
  ex_full = np.repeat(ex[..., None], pay.shape[-1], axis=-1)
  grid = np.arange(time_step, unique_ex[-1], time_step)
  sim_times = np.unique(np.concatenate([unique_ex, grid]))
  tau = pay - ex_full
  curve_times = np.unique(tau.reshape(-1))
  return dict(batch_shape=batch_shape, unique_ex=unique_ex,
              ex_index=ex_index.reshape(ex.shape), ex_full=ex_full,
              sim_times=sim_times, tau=tau, curve_times=curve_times,
              n_ex=ex.shape[-1])


def _exercise_bonds(model, plan, r_t, pay):
  """P(t_e, T_j) [N, batch, n_ex, n_pay] from the short rate at each
  exercise date: the entries of the reference's discount-curve tensor
  (swaption.py:655-672) that the payoffs read, without materialising the
  whole [N, tenors, times] curve."""
  dev = r_t.device
  # the short rate once per exercise date [N, batch, n_ex], and the bond
  # as exp(C - r G) with the host constant C = ln(P(0,T)/P(0,t)) + f(0,t) G
  # - y G^2 / 2: one fused multiply-add and one exponential over the
  # [N, batch, n_ex, n_pay] tensor (the same bonds as model._bond_prices
  # up to rounding)
  ex = plan['ex_full'][..., 0]
  sim_idx = np.searchsorted(plan['sim_times'], ex)
  r_e = r_t[:, torch.as_tensor(sim_idx, device=dev), 0]
  f0, p0big, g, y = model._bond_constants(plan['ex_full'], pay)
  c = np.log(p0big) + f0 * g - 0.5 * y * g * g
  as_t = lambda a: torch.as_tensor(a, dtype=r_t.dtype, device=dev)
  return torch.addcmul(as_t(c), r_e[..., None], as_t(-g)).exp_()


def bonds_from_curve_paths(plan, p_t_tau, pay):
  """The same bonds gathered from a full sample_discount_curve_paths
  tensor [N, tenors, times, 1]."""
  dev = p_t_tau.device
  n = p_t_tau.shape[0]
  sim_idx = np.searchsorted(plan['sim_times'], plan['ex_full'].reshape(-1))
  cur_idx = np.searchsorted(plan['curve_times'], plan['tau'].reshape(-1))
  b = p_t_tau[:, torch.as_tensor(cur_idx, device=dev),
              torch.as_tensor(sim_idx, device=dev), 0]
  return b.reshape([n] + list(pay.shape))


def _lsm_inputs(plan, bonds, r_t, pay, dcf, cpn, tdt):
  """(short rates [N, n_u, 1] at the union of exercise dates, payoffs
  [N, B, n_u], path discount factors [N, n_u]) from the exercise-date bonds
  [N, batch, n_ex, n_pay] and the short rates [N, k, 1]."""
  dev = r_t.device
  n = r_t.shape[0]
  sim_times = plan['sim_times']
  dt = np.concatenate([[0.0], sim_times[1:] - sim_times[:-1]])
  unique_ex = plan['unique_ex']
  u_idx = np.searchsorted(sim_times, unique_ex)
  # path discount factors at the exercise dates only, exp(-sum r dt) up to
  # each date: one [N, k] x [k, n_u] product with the masked step lengths
  # (the same values as the cumulative product of the per-step factors, up
  # to rounding, without scanning all k dates of every path)
  steps = np.where(np.arange(dt.shape[0])[:, None] <= u_idx[None, :],
                   dt[:, None], 0.0)
  disc_u = torch.exp(-(r_t[..., 0] @ torch.as_tensor(steps, dtype=tdt,
                                                      device=dev)))
  # swap = (1 - P_last) - sum_j c_j P_j = 1 - sum_j w_j P_j with w_j = c_j
  # plus 1 at the last payment: one float64 product of the bonds
  # [N, batch * n_ex * n_pay] with the block-diagonal weights
  n_pay = bonds.shape[-1]
  n_k = int(np.prod(bonds.shape[1:-1]))
  w = np.zeros((n_k * n_pay, n_k))
  cw = np.broadcast_to(cpn * dcf, bonds.shape[1:]).reshape(n_k, n_pay).copy()
  cw[:, -1] += 1.0
  for k in range(n_k):
    w[k * n_pay:(k + 1) * n_pay, k] = cw[k]
  swap = (1.0 - bonds.reshape(n, n_k * n_pay) @ torch.as_tensor(
      w, dtype=tdt, device=dev)).reshape(bonds.shape[:-1])  # [N, batch, n_ex]
  # payoffs on the union of exercise dates; a date that is not one of a
  # swaption's own dates pays nothing for it (swaption.py:876-935)
  n_u = unique_ex.shape[0]
  shape = plan['batch_shape']
  flat_b = int(np.prod(shape)) if shape else 1
  n_ex = plan['n_ex']
  swap_f = swap.reshape(n, flat_b, n_ex)
  exi = torch.as_tensor(plan['ex_index'].reshape(flat_b, n_ex), device=dev)
  pay_u = torch.zeros((n, flat_b, n_u), dtype=tdt, device=dev)
  is_ex = torch.zeros((flat_b, n_u), dtype=torch.bool, device=dev)
  for b in range(flat_b):
    pay_u[:, b, exi[b]] = swap_f[:, b]
    is_ex[b, exi[b]] = True
  payoff_table = torch.where(is_ex[None], torch.relu(pay_u),
                             torch.zeros_like(pay_u))
  return (r_t[:, torch.as_tensor(u_idx, device=dev)], payoff_table,
          disc_u)


def _price_on_paths(plan, bonds, r_t, pay, dcf, cpn, basis, tdt):
  short_rate, payoff_table, dfs = _lsm_inputs(plan, bonds, r_t, pay, dcf,
                                              cpn, tdt)
  # the table payoff and per-path discount factors keep the regression in
  # the native LSM kernels
  payoff_fn = lsm.make_table_payoff(payoff_table)
  return lsm.least_square_mc(short_rate, list(range(payoff_table.shape[-1])),
                             payoff_fn, basis,
                             discount_factors=dfs[:, None, :], dtype=tdt)
