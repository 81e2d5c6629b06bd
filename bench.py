"""Benchmark: fused FP32 Heston full-truncation log-Euler on one MI355X.

Workload (the headline configuration): S0=100, v0=0.04, kappa=2, theta=0.04,
xi=0.3, rho=-0.7, r=0.03, T=1, 2^20 Sobol paths (730 dimensions, Joe-Kuo
direction numbers, 32-digit digital shift drawn from the engine's Philox
stream with seed 7), 365 steps on the reference grid of
euler_sampling.sample(num_time_steps=365, times=maturities), payoffs fused
for 64 strikes linspace(60, 140) x 4 maturities {0.25, 0.5, 0.75, 1.0}.
One step = one full pricing run (simulation + payoff reduction + finalize)
with inputs already resident on the GPU.

  python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'tf-quant-finance_amd'))
sys.path.insert(0, ROOT)

METRIC = ('Heston path-steps/s on 1 GPU; options+Greeks/s at 1e-3 rel err vs '
          'FP64')
FLOP_PER_PATH_STEP = 73  # algorithmic count per path-step, FP32
# HBM bytes per launch of the headline kernel at 2^20 paths: FETCH_SIZE and
# WRITE_SIZE of two rocprofv3 --pmc passes, each scaled by the same counter
# on a 256 MiB device copy (profiles/euler_heston_ft_sobol_pmc.md).
TRAFFIC_BYTES_2POW20 = 23.87e6
PEAK_FP32 = 157.3e12
PEAK_HBM = 8.0e12

PARAMS = dict(s0=100.0, v0=0.04, kappa=2.0, theta=0.04, xi=0.3, rho=-0.7,
              r=0.03)
MATURITIES = (0.25, 0.5, 0.75, 1.0)
NUM_STEPS = 365


# ------------------------------------------------------------ CPU legs ----
# The reference's TF CPU path cannot run here, so each CPU baseline times the
# oracle's numpy restatement of the same algorithm on this machine's host
# cores: one process per core (numpy's elementwise work is single-threaded;
# the workers are started with `spawn` and import numpy, not torch), path
# chunks spread over them, one warm-up run, then the median of three.


def cpu_cores():
  """Host cores this process may use: its CPU affinity, capped by
  OMP_NUM_THREADS when that is set (the machine's CPU share)."""
  n = len(os.sched_getaffinity(0))
  cap = os.environ.get('OMP_NUM_THREADS', '')
  if cap.isdigit() and int(cap) > 0:
    n = min(n, int(cap))
  return max(n, 1)


def _worker_init():
  for p in (ROOT, os.path.join(ROOT, 'tf-quant-finance_amd')):
    if p not in sys.path:
      sys.path.insert(0, p)
  try:
    import threadpoolctl
    threadpoolctl.threadpool_limits(1)
  except ImportError:
    pass


class CpuPool:
  """`cores` spawned worker processes (context manager)."""

  def __init__(self, cores):
    import multiprocessing as mp
    self.cores = cores
    self.pool = mp.get_context('spawn').Pool(cores, initializer=_worker_init)

  def __enter__(self):
    return self

  def __exit__(self, *exc):
    self.pool.terminate()
    self.pool.join()

  def map(self, fn, items):
    return self.pool.map(fn, items, chunksize=1)


def median_time(run, runs=3):
  """One warm-up call, then the median wall time of `runs` calls."""
  run()
  ts = []
  for _ in range(runs):
    t0 = time.perf_counter()
    run()
    ts.append(time.perf_counter() - t0)
  return float(np.median(ts)), ts


def _w_heston_sobol(a):
  """Headline scheme on one chunk of paths: the oracle's 32-digit shifted
  Sobol points, the engine's normal transform, the FT step loop and the
  payoffs (float32 state)."""
  from oracle import heston as oheston
  from oracle import normal as onormal
  first, m, shift, all_times, keep = a
  steps = all_times.shape[0] - 1
  ints = oheston.shifted_sobol_ints(2 * steps, m, shift, first=first)
  z = onormal.shifted_sobol_normals(ints).reshape(m, steps, 2)
  oheston.ft_euler_calls(np.log(PARAMS['s0']), PARAMS['v0'], PARAMS['kappa'],
                         PARAMS['theta'], PARAMS['xi'], PARAMS['rho'],
                         PARAMS['r'], np.linspace(60, 140, 64), all_times,
                         keep, z.astype(np.float32), dtype=np.float32)
  return m


def cpu_baseline_headline(pool, all_times, keep, shift, n_paths,
                          chunk=2**14):
  steps = all_times.shape[0] - 1
  work = [(f, min(chunk, n_paths - f), shift, all_times, keep)
          for f in range(0, n_paths, chunk)]
  t, ts = median_time(lambda: pool.map(_w_heston_sobol, work))
  return {'value': n_paths * steps / t, 'unit': 'path-steps/s',
          'cores': pool.cores, 'kind': 'port',
          'sample': '%d paths x %d steps%s, numpy FP32 restatement of the '
                    'same scheme incl. the Sobol points and normals, %d '
                    'processes, median of 3 after a warm-up (%.2f s; runs %s '
                    's)' % (n_paths, steps,
                            ' (the full workload)' if n_paths == 2**20 else '',
                            pool.cores, t, [round(x, 2) for x in ts])}


def _w_gbm_euler(a):
  """Plain GBM Euler on S, float64 (euler_sampling's step loop with the
  GeometricBrownianMotion coefficients), 8 discounted calls."""
  from oracle import euler as oeuler
  seed, m, all_times, strikes = a
  steps = all_times.shape[0] - 1
  z = np.random.default_rng(seed).standard_normal((steps, m, 1))
  drift, vol = oeuler.gbm_fns(0.05, 0.2)
  keep = np.zeros(steps + 1, bool)
  keep[-1] = True
  x = oeuler.run(drift, vol, all_times, keep, 1, np.full((m, 1), 100.0), z,
                 np.float64)
  return np.exp(-0.05) * np.maximum(x[:, 0, :] - strikes, 0).sum(0)


def cpu_baseline_gbm_euler(pool, all_times, n_paths=100000, chunk=6250):
  work = [(i, min(chunk, n_paths - f), all_times, np.linspace(80, 120, 8))
          for i, f in enumerate(range(0, n_paths, chunk))]
  t, _ = median_time(lambda: pool.map(_w_gbm_euler, work))
  steps = all_times.shape[0] - 1
  return {'value': n_paths * steps / t, 'unit': 'path-steps/s',
          'cores': pool.cores, 'kind': 'port',
          'sample': 'full size (%d paths x %d steps, float64, numpy normals), '
                    '%d processes, median of 3 after a warm-up (%.3f s)' % (
                        n_paths, steps, pool.cores, t)}


def _w_contract(a):
  """One contract of the heterogeneous batch: price and pathwise Greeks
  (GBM log-Euler: delta, vega, rho; Heston full truncation: forward
  tangents of the step loop for delta, v0, r, kappa, theta, xi, rho)."""
  from oracle import heston as oheston
  c, m, seed = a
  n, T, r = c['n_steps'], c['maturity'], c['rate']
  rng = np.random.default_rng(seed)
  if c['kind'] == 0:
    dt = T / n
    sig = c['sigma']
    x = np.full(m, np.log(c['s0']))
    w = np.zeros(m)
    for _ in range(n):
      dw = rng.standard_normal(m) * np.sqrt(dt)
      x = x + (r - 0.5 * sig * sig) * dt + sig * dw
      w = w + dw
    s = np.exp(x)
    disc = np.exp(-r * T)
    itm = s > c['strike']
    pay = disc * np.where(itm, s - c['strike'], 0)
    sb = disc * np.where(itm, s, 0)
    g = np.stack([pay, sb / c['s0'], sb * (w - sig * T), -T * pay + sb * T])
    return g.mean(1)
  z = rng.standard_normal((m, n, 2))
  mean, _ = oheston.ft_euler_call_greeks(
      np.log(c['s0']), c['v0'], c['kappa'], c['theta'], c['xi'], c['rho'], r,
      c['strike'], np.linspace(0.0, T, n + 1), z)
  return mean


def cpu_baseline_batch(pool, contracts, total_path_steps, n_sample=64,
                       max_paths=2048):
  """64 contracts, every 64th of the table (stratified by position), at
  most `max_paths` paths each; the time is extrapolated linearly in
  path-steps to the whole table."""
  idx = list(range(0, len(contracts), max(len(contracts) // n_sample, 1)))
  idx = idx[:n_sample]
  work = [(contracts[i], min(contracts[i]['n_paths'], max_paths), i)
          for i in idx]
  sample_ps = sum(m * c['n_steps'] for c, m, _ in work)
  t, _ = median_time(lambda: pool.map(_w_contract, work))
  t_full = t * total_path_steps / sample_ps
  return {'value': len(contracts) / t_full, 'unit': 'options+Greeks/s',
          'cores': pool.cores, 'kind': 'port',
          'sample': '%d contracts (every %d-th), at most %d paths each '
                    '(%.3g path-steps) in %.2f s (median of 3 after a '
                    'warm-up, %d processes); extrapolated linearly in '
                    'path-steps to the %.3g of the whole table' % (
                        len(work), len(contracts) // n_sample, max_paths,
                        sample_ps, t, pool.cores, total_path_steps)}


def _w_basket(a):
  from oracle import gbm as ogbm
  s0, sig, corr, m, steps, seed = a
  z = np.random.default_rng(seed).standard_normal((m, steps, len(s0)))
  p = ogbm.basket_call_paths(s0, sig, corr, 0.02, 1.0, 100.0, z)
  return p['pay'].sum()


def cpu_baseline_basket(pool, s0, sig, corr, n_paths=262144, steps=64,
                        chunk=8192):
  work = [(s0, sig, corr, min(chunk, n_paths - f), steps, i)
          for i, f in enumerate(range(0, n_paths, chunk))]
  t, _ = median_time(lambda: pool.map(_w_basket, work))
  return {'value': n_paths * steps / t, 'unit': 'path-steps/s',
          'cores': pool.cores, 'kind': 'port',
          'sample': 'full size (%d paths x %d steps x %d assets, float64, '
                    'price + 33 pathwise Greeks), %d processes, median of 3 '
                    'after a warm-up (%.2f s)' % (n_paths, steps, len(s0),
                                                   pool.cores, t)}


def _w_lsm(a):
  """Polynomial LSM of the oracle on a subsample: the standardised basis
  (the engine's judge) and the reference's raw basis with pinv."""
  from oracle import lsm as olsm
  paths, ex, df, n_cal, cores, standardise = a
  try:
    import threadpoolctl
    threadpoolctl.threadpool_limits(cores)
  except ImportError:
    pass
  return float(olsm.least_square_mc(
      paths, ex, olsm.make_basket_call_max_payoff([100.0]),
      olsm.make_polynomial_basis(2, standardise), discount_factors=df,
      num_calibration_samples=n_cal)[0])


def accuracy_check(prices, ses):
  """FP32 2^20-path prices against the engine's float64 run of the same
  scheme at 2^24 paths (the float64 path is pinned to the numpy oracle at
  1e-12 by the parity tests, and to the semi-analytic Heston price by
  tests/test_gpu_core.py)."""
  import torch
  from tqf.engine import pricing
  ref = pricing.heston_ft_calls(num_paths=2**24, num_time_steps=NUM_STEPS,
                                maturities=MATURITIES, grid='reference',
                                random='sobol_shifted', seed=7,
                                dtype=torch.float64, **PARAMS)
  rp = ref['prices'].cpu().numpy()
  rse = ref['stderrs'].cpu().numpy()
  big = rp > 1.0
  err = np.abs(prices - rp)
  comb = np.sqrt(ses**2 + rse**2)
  return {
      'vs': 'engine float64, same scheme, 2^24 paths',
      'max_rel_err_price_gt_1': float(np.max(err[big] / rp[big])),
      'max_abs_err': float(np.max(err)),
      'max_err_in_combined_se': float(np.max(err / comb)),
      'frac_within_3se': float(np.mean(err <= 3 * comb)),
      'max_stderr': float(np.max(ses)),
  }


def batch_accuracy(contracts, batch_prices, seed=11, every=256):
  """Batched FP32 prices against a float64 run of the same scheme on the
  same Philox streams (the plain fused kernel, per contract) for every
  `every`-th contract: the FP32 arithmetic error, free of Monte Carlo noise.
  """
  import torch
  from tqf.engine import euler as engine
  dev = torch.device('cuda', 0)
  errs = []
  for i in range(0, len(contracts), every):
    c = contracts[i]
    n, T, r = c['n_steps'], c['maturity'], c['rate']
    all_times = np.linspace(0.0, T, n + 1)
    keep = np.zeros(n + 1, bool)
    keep[-1] = True
    rng = engine.RngSpec('philox', seed=seed, stream_id=c.get('stream_id', i))
    if c['kind'] == 0:
      spec = ('const', r - 0.5 * c['sigma']**2, c['sigma'])
      x0 = [np.log(c['s0'])]
    else:
      spec = ('heston', r, c['kappa'], c['theta'], c['xi'], c['rho'], 0.0)
      x0 = [np.log(c['s0']), c['v0']]
    res = engine.run(spec, all_times, keep, 1, x0, c['n_paths'], rng,
                     torch.float64, mode='payoff', strikes=[c['strike']],
                     discount=[np.exp(-r * T)], dev=dev)
    ref = float(res['prices'].cpu()[0, 0])
    errs.append(abs(float(batch_prices[i]) - ref) / max(abs(ref), 1e-2))
  return {'vs': 'float64 plain kernel, same Philox streams, %d contracts'
                % len(errs),
          'max_rel_err_price': float(np.max(errs))}


def contract_loop(contracts, seed=11, h=1e-3):
  """Returns run(greeks) -> seconds for pricing `contracts` one tqf_euler
  call each (uniform grid, Philox sub-stream per contract, payoff mode), the
  loop a caller without the batched kernel would write.  With greeks=True
  each Greek of the batched kernel is added by central differences with
  common random numbers (relative bump h).  Every launch is prepared before
  the timed loop."""
  import ctypes
  import torch
  from tqf import _native
  from tqf.engine import euler as engine
  dev = torch.device('cuda', 0)
  base, bumped = [], []
  for i, c in enumerate(contracts):
    n, T, r = c['n_steps'], c['maturity'], c['rate']
    all_times = np.linspace(0.0, T, n + 1).astype(np.float32)
    keep = np.zeros(n + 1, bool)
    keep[-1] = True
    rng = engine.RngSpec('philox', seed=seed, stream_id=c.get('stream_id', i))
    if c['kind'] == 0:
      spec = ('const', r - 0.5 * c['sigma']**2, c['sigma'])
      x0 = [np.log(c['s0'])]
    else:
      spec = ('heston', r, c['kappa'], c['theta'], c['xi'], c['rho'], 0.0)
      x0 = [np.log(c['s0']), c['v0']]
    L = engine.prepare(spec, all_times, keep, 1, x0, c['n_paths'], rng,
                       torch.float32, mode='payoff', strikes=[c['strike']],
                       discount=[np.exp(-r * T)], dev=dev)
    base.append(L)
    # bumped variants share the grid, stream and workspace; each has its
    # own model parameters or initial state and its own outputs
    variants = []
    if c['kind'] == 0:
      sig = c['sigma']
      bumps = [('x0', 0), ('p', ('sigma',)), ('p', ('r',))]
    else:
      bumps = [('x0', 0), ('x0', 1), ('p', ('r',)), ('p', (1,)), ('p', (2,)),
               ('p', (3,)), ('p', (4,))]
    for kind, what in bumps:
      for sgn in (1.0, -1.0):
        m = _native.Model.from_buffer_copy(L.m)
        x0_t = L.x0_t
        if kind == 'x0':
          x = np.asarray(x0, np.float64).copy()
          if what == 0:
            x[0] += np.log1p(sgn * h)
          else:
            x[1] *= 1 + sgn * h
          x0_t = torch.as_tensor(x, dtype=torch.float32, device=dev)
        elif c['kind'] == 0:
          sig2 = c['sigma'] * (1 + sgn * h) if what == ('sigma',) else c['sigma']
          r2 = r * (1 + sgn * h) if what == ('r',) else r
          m.p[0], m.p[1] = r2 - 0.5 * sig2 * sig2, sig2
        else:
          k = 0 if what == ('r',) else what[0]
          m.p[k] = m.p[k] * (1 + sgn * h)
        o = _native.Output.from_buffer_copy(L.o)
        outs = torch.empty(4, dtype=torch.float64, device=dev)
        o.prices, o.stderrs, o.sums = (outs.data_ptr(), outs.data_ptr() + 8,
                                       0)
        variants.append((m, o, x0_t, outs))
    bumped.append(variants)

  def call(L, m, o, x0_t):
    _native.call('tqf_euler', ctypes.byref(m), ctypes.byref(L.g),
                 ctypes.byref(L.r), ctypes.byref(o),
                 ctypes.c_int64(L.num_samples),
                 ctypes.c_void_p(x0_t.data_ptr()),
                 ctypes.c_int64(L.x0_stride), ctypes.c_int32(0),
                 _native.stream_ptr(dev))

  def run(greeks):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for L, variants in zip(base, bumped):
      L.launch()
      if greeks:
        for m, o, x0_t, _ in variants:
          call(L, m, o, x0_t)
    torch.cuda.synchronize()
    return time.perf_counter() - t0
  run(False)  # warm-up
  return run


SECONDARY = ('gbm_euler_f64', 'batch_4096', 'learned_lsm_5asset',
             'hw_bermudan_2x2e20', 'basket16_greeks', 'paths_and_points')
HBM_PEAK = 8e12  # bytes/s


def secondary_metrics(only=None, pool=None):
  """One timed run each of the other configurations (not the headline), or
  of the one named by `only`; with a CpuPool each also times its CPU
  baseline."""
  import torch

  def want(name):
    return only is None or only == name
  from tqf.engine import batch as batch_lib
  from tqf.engine import greeks
  from tqf.engine import euler as engine
  from tqf.engine import pricing
  out = {}

  def timed(fn, reps=3):
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
      r = fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / 1e3 / reps, r

  # plain method: GBM Euler float64, 1e5 paths x 100 steps, 8 calls
  if want('gbm_euler_f64'):
    _gbm_euler(out, timed, pricing, engine, pool)
  if want('batch_4096'):
    _batch(out, timed, batch_lib, pool)
  if want('learned_lsm_5asset'):
    _learned_lsm(out, pool)
  if want('hw_bermudan_2x2e20'):
    _hw_bermudan(out)
  if want('basket16_greeks'):
    _basket16(out, timed, greeks, pool)
  if want('paths_and_points'):
    _paths_and_points(out, timed, engine)
  return out


FLOP_GBM_EULER = 8.5  # FP64 per path-step (update 3, half a Box-Muller
                      # pair 5, payoff 0.4 amortised)
PEAK_FP64 = 78.6e12


def _gbm_euler(out, timed, pricing, engine, pool):
  import torch
  all_times, keep = pricing.heston_grid_for([1.0], 100, 'reference',
                                            torch.float64)
  c1 = pricing.european_calls(('gbm', 0.05, 0.2), [100.0],
                              np.linspace(80, 120, 8), all_times, keep,
                              100000, engine.RngSpec('philox', seed=42),
                              discount_rate=0.05, dtype=torch.float64,
                              prepare=True)
  t, _ = timed(c1.launch, 10)
  ach = FLOP_GBM_EULER * 1e7 / t
  out['gbm_euler_f64'] = {
      'value': 1e7 / t, 'unit': 'path-steps/s', 'seconds': t,
      'roofline': {'bound': 'valu', 'achieved': ach / 1e12,
                   'peak': PEAK_FP64 / 1e12, 'unit': 'TFLOP/s',
                   'frac': ach / PEAK_FP64,
                   'flop_per_path_step': FLOP_GBM_EULER,
                   'note': 'one launch of 1e5 paths fills 391 workgroups: '
                           'latency-bound, no roofline target'}}
  if pool is not None:
    out['gbm_euler_f64']['cpu_baseline'] = cpu_baseline_gbm_euler(
        pool, np.asarray(all_times, np.float64))


def _batch(out, timed, batch_lib, pool):
  # batched heterogeneous contracts with Greeks
  cs = batch_lib.random_contracts(4096)
  cfg, tune = batch_lib.autotune(cs, cache_path=None)
  plan = batch_lib.BatchPlan(cs, **cfg)
  g = plan.capture()
  t, _ = timed(lambda: g.replay(), 2)
  acc = batch_accuracy(cs, plan.out[:, 0].cpu().numpy())
  # the other Greek mode on the whole table, timed: the checkpointed
  # reverse sweep and the forward tangents are two implementations of the
  # same pathwise estimator on the same draws, so they must agree for every
  # contract and Greek up to float32 rounding
  other = 'adjoint' if cfg['greeks'] == 'tangent' else 'tangent'
  plan2 = batch_lib.BatchPlan(cs, tile=cfg['tile'], n_blocks=cfg['n_blocks'],
                              greeks=other)
  t2, _ = timed(plan2.launch, 2)
  a = plan.out.cpu().numpy()
  b = plan2.out.cpu().numpy()
  means = a[:, :8]
  scale = np.maximum(np.abs(means), 1e-3 * np.abs(a[:, :1]))
  fin = np.isfinite(means) & np.isfinite(b[:, :8])
  # Feller-violating Heston contracts: the variance reaches 0 on some
  # paths, where the pathwise variance-parameter derivatives carry
  # 1/sqrt(v) factors that can overflow float32
  feller = np.array([c['kind'] == batch_lib.KIND_GBM or
                     2 * c['kappa'] * c['theta'] >= c['xi']**2 for c in cs])
  # relative differences where both are finite and the scale is positive
  # (a zero price and Greek, e.g. a contract with no path in the money,
  # is compared absolutely)
  ok = fin & (scale > 0)
  rel = np.zeros_like(means)
  rel[ok] = np.abs(b[:, :8][ok] - means[ok]) / scale[ok]
  zero = fin & (scale == 0)
  acc['greeks_%s_vs_%s' % (other, cfg['greeks'])] = {
      'contracts': len(cs),
      'entries_compared': int(ok.sum()),
      'max_rel_diff': float(rel.max()),
      'max_rel_diff_feller_contracts': float(rel[feller].max()),
      'max_abs_diff_zero_scale': float(np.max(
          np.abs(b[:, :8][zero] - means[zero]), initial=0.0)),
      'contracts_with_nonfinite_greeks': {
          cfg['greeks']: int(np.sum(~np.all(np.isfinite(means), 1))),
          other: int(np.sum(~np.all(np.isfinite(b[:, :8]), 1)))},
      'prices_identical': bool(np.array_equal(a[:, 0], b[:, 0]))}
  out['batch_4096'] = {'value': 4096 / t, 'unit': 'options+Greeks/s',
                       ('%s_seconds' % other): t2,
                       'seconds': t, 'check': acc,
                       'path_steps': plan.total_path_steps,
                       'path_steps_per_s': plan.total_path_steps / t,
                       'flop': plan.total_flop,
                       'roofline': {
                           'bound': 'valu',
                           'achieved': plan.total_flop / t / 1e12,
                           'peak': PEAK_FP32 / 1e12, 'unit': 'TFLOP/s',
                           'frac': plan.total_flop / t / PEAK_FP32,
                           'kernel': 'batch_kernel',
                           'flop_per_path_step': 'GBM 10, Heston 95 '
                                                 '(tangent) / 99 (adjoint)'},
                       'autotuned': cfg,
                       'candidates_s': {'%s/%d/%d' % (r['greeks'], r['tile'],
                                                      r['n_blocks']):
                                        round(r['seconds'], 4) for r in tune}}
  # the same contracts through the plain fused kernel, one call per contract
  # (price only, and price + central-difference Greeks with common random
  # numbers: 2 more calls per Greek)
  loop = contract_loop(cs)
  t_price = loop(False)
  t_greeks = loop(True)
  out['batch_4096']['loop_price_only_seconds'] = t_price
  out['batch_4096']['loop_fd_greeks_seconds'] = t_greeks
  out['batch_4096']['speedup_vs_loop_fd_greeks'] = t_greeks / t
  out['batch_4096']['speedup_vs_loop_price_only'] = t_price / t
  if pool is not None:
    out['batch_4096']['cpu_baseline'] = cpu_baseline_batch(
        pool, cs, plan.total_path_steps)


FLOP_MLP_SAMPLE_STEP = 27264  # forward 4544 MAC + backward 2x, per sample
FLOP_LSM_DATE = 6.4e10  # one K = 243 regression date at 2^20 calibration
                        # paths as a dense Gram (Gram, X^T y, continuation)
# what the engine executes per date instead: the 25 x 152 box moments of the
# degree-2 tensor basis of 5 assets (2 flop per moment and path; the MFMA
# tiles pad them to 32 x 160)
FLOP_LSM_MOMENT_DATE = 2.0 * 25 * 152 * 2**20
PEAK_BF16 = 2.5e15


def _learned_lsm(out, pool):
  import torch
  # learned LSM at full size: 5-asset Bermudan max-call, 2^20 training +
  # 2^20 pricing paths, 9 dates x 20 steps, MLP vs degree-2 polynomial
  from tqf.engine import mlp_lsm
  from tqf.models import longstaff_schwartz as lsm_lib
  ex = list(range(1, 10))
  n_cal = 2**20
  # one untimed call first: it captures the per-date training graph that
  # every later call of this size replays (about 0.1 s, once per process)
  paths, df = mlp_lsm.bermudan_max_call_paths(2 * n_cal)
  mlp_lsm.mlp_lsm_price(paths, ex, 100.0, df, n_cal, judge=False)
  del paths
  # the whole pipeline (paths, learned policy, polynomial judge, paired
  # statistic): median of three runs, and the path simulation alone
  totals, sims = [], []
  for _ in range(3):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    paths, df = mlp_lsm.bermudan_max_call_paths(2 * n_cal)
    torch.cuda.synchronize()
    sims.append(time.perf_counter() - t0)
    res = mlp_lsm.mlp_lsm_price(paths, ex, 100.0, df, n_cal)
    torch.cuda.synchronize()
    totals.append(time.perf_counter() - t0)
  total = float(np.median(totals))
  dfp = lsm_lib._padded_discounts(df, len(ex), torch.float64)
  payoff = (lsm_lib.PAYOFF_MAX_CALL, np.array([100.0]))

  def wall(fn):
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t1
  # the two parts on their own (each already ran once above): the
  # polynomial judge and the MLP training + exercise policy
  t_poly = wall(lambda: lsm_lib._native_lsm(paths, ex, payoff, 2, dfp, n_cal,
                                            torch.float64, paths.device))
  t_mlp = wall(lambda: mlp_lsm.mlp_lsm_price(paths, ex, 100.0, df, n_cal,
                                             judge=False))
  mlp_flop = FLOP_MLP_SAMPLE_STEP * 65536 * 200 * (len(ex) - 1)
  reg_flop = FLOP_LSM_DATE * (len(ex) - 1)
  mom_flop = FLOP_LSM_MOMENT_DATE * (len(ex) - 1)
  line = {
      'seconds': total,
      'seconds_runs': [round(x, 4) for x in totals],
      'paths_seconds': float(np.median(sims)),
      'mlp_price': res['mlp_price'], 'mlp_se': res['mlp_se'],
      'poly_price': res['poly_price'], 'poly_se': res['poly_se'],
      'gain': res['gain'], 'gain_se_paired': res['gain_se'],
      'gain_se_unpaired': res['gain_se_unpaired'],
      'accepted': res['accepted'], 'price': res['price'],
      'poly_seconds': t_poly, 'mlp_seconds': t_mlp,
      'roofline': {'bound': 'mfma', 'achieved': mlp_flop / t_mlp / 1e12,
                   'peak': PEAK_BF16 / 1e12, 'unit': 'TFLOP/s',
                   'frac': mlp_flop / t_mlp / PEAK_BF16,
                   'kernel': 'mlp training (mlp_grad_kernel + '
                             'mlp_reduce_adam_kernel), 1600 Adam steps of '
                             '65,536 samples',
                   'flop_per_sample_step': FLOP_MLP_SAMPLE_STEP},
      'roofline_regression': {
          'bound': 'mfma', 'achieved': mom_flop / t_poly / 1e12,
          'peak': PEAK_FP64 / 1e12, 'unit': 'TFLOP/s',
          'frac': mom_flop / t_poly / PEAK_FP64,
          'flop_per_date': FLOP_LSM_MOMENT_DATE,
          'basis': 'float64 moment product (25 x 152 moments per path, '
                   '2^20 calibration paths) per regression date over the '
                   'wall time of the whole polynomial LSM (moments, '
                   'Cholesky, exercise updates)',
          'dense_gram_equivalent_tflops': reg_flop / t_poly / 1e12}}
  if pool is not None:
    # the oracle's polynomial LSM on a subsample of the same paths: the
    # first 2^14 calibration rows and the first 2^14 pricing rows
    m = 2**14
    sub = torch.cat([paths[:m], paths[n_cal:n_cal + m]]).contiguous()
    sub_np = sub.cpu().numpy()
    gpu_sub = float(lsm_lib._native_lsm(sub, ex, payoff, 2, dfp, m,
                                        torch.float64, paths.device)[0])
    item = [(sub_np, ex, df, m, pool.cores, True)]
    t_cpu, ts = median_time(lambda: pool.map(_w_lsm, item))
    want = pool.map(_w_lsm, item)[0]
    raw = pool.map(_w_lsm, [(sub_np, ex, df, m, pool.cores, False)])[0]
    scale = 2 * n_cal / (2 * m)
    line['cpu_baseline'] = {
        'value': t_cpu * scale, 'unit': 's (polynomial LSM, extrapolated)',
        'cores': pool.cores, 'kind': 'port',
        'sample': 'polynomial LSM only (K = 243, 8 regressions) on %d + %d '
                  'rows of the same paths in %.2f s (median of 3 after a '
                  'warm-up, BLAS on %d threads), extrapolated linearly in '
                  'rows to 2^20 + 2^20 (x%d); the engine takes %.3f s for '
                  'the same at full size' % (m, m, t_cpu, pool.cores, scale,
                                             t_poly)}
    line['check'] = {
        'vs': 'oracle polynomial LSM (standardised basis, numpy pinv) on '
              'the %d + %d-row subsample' % (m, m),
        'engine_price': gpu_sub, 'oracle_price': want,
        'rel_err': abs(gpu_sub - want) / abs(want),
        'reference_semantics_price': raw,
        'reference_semantics': 'pinv of the raw 243-column Gram with TF\'s '
                               'cutoff (lsm.py:369-383) on the same rows'}
  out['learned_lsm_5asset'] = line
  del paths


def _hw_bermudan(out):
  import torch
  # Hull-White Bermudan swaptions (5nc1 and 5nc2, semiannual, a = 3%,
  # sigma = 1%, flat 1% curve), 2^20 paths: short rate in the fused kernel,
  # regression in the native LSM with per-path discount factors
  from tqf.models.hull_white import bermudan_swaption_price
  from tqf.math.random.multivariate_normal import RandomType
  base = [1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0]
  st = np.array([[[base[min(i + j, 8)] for j in range(8)] for i in range(8)],
                 [[base[min(i + j, 8)] for j in range(8)]
                  for i in range(2, 10)]])
  end = np.clip(st + 0.5, 0.0, 5.0)
  ex = np.array([base[:8], base[2:] + [5.0]])

  def berm():
    return bermudan_swaption_price(
        exercise_times=ex, floating_leg_start_times=st,
        floating_leg_end_times=end, fixed_leg_payment_times=end,
        floating_leg_daycount_fractions=end - st,
        fixed_leg_daycount_fractions=end - st,
        fixed_leg_coupon=0.011 * np.ones_like(end),
        reference_rate_fn=lambda t: 0.01 * torch.ones_like(
            torch.as_tensor(t)), notional=100.0, mean_reversion=0.03,
        volatility=0.01, num_samples=2**20, time_step=0.1,
        random_type=RandomType.STATELESS_ANTITHETIC, seed=[0, 0],
        dtype=torch.float64)
  berm()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  prices = berm().cpu().numpy()
  t = time.perf_counter() - t0
  out['hw_bermudan_2x2e20'] = {'seconds': t, 'prices': prices.tolist(),
                               'reference_values': [1.8892, 1.6633]}


def _paths_and_points(out, timed, engine):
  """The other call surfaces at full size, one warm call then timed
  calls: the C2 scheme in path-store mode (every step of (log S, v) written
  coalesced [step][d][N], HBM-bound), sobol.sample of the C2 draw set,
  HestonModel (QE) and GeometricBrownianMotion (exact) sample_paths,
  Halton points and the pricing platform's American options."""
  import torch
  import tqf
  from tqf.engine import pricing
  from tqf.experimental import american_option
  from tqf.math.random import halton
  from tqf.math.random import sobol as tsobol
  res = {}
  n = 2**20
  # path store: 2^20 paths x 365 steps, all 366 records of (log S, v) FP32
  all_times = np.arange(366) / 365.0
  keep = np.ones(366, bool)
  rng = engine.RngSpec('sobol_shifted',
                       shift=pricing.sobol_digital_shift(730, 7),
                       first_index=0)
  lp = engine.prepare(('heston', 0.03, 2.0, 0.04, 0.3, -0.7, 0.0), all_times,
                      keep, 366, np.array([np.log(100.0), 0.04]), n, rng,
                      torch.float32, mode='store')
  lp.launch()
  t, _ = timed(lp.launch, 3)
  nbytes = 366 * 2 * n * 4
  res['pathstore_heston_ft'] = {
      'value': nbytes / t / 1e9, 'unit': 'GB/s', 'seconds': t,
      'bytes_written': nbytes, 'frac_hbm_peak': nbytes / t / HBM_PEAK,
      'path_steps_per_s': n * 365 / t}
  del lp
  torch.cuda.empty_cache()
  # the same in the reference layout [N, k, d] (what euler_sampling.sample
  # returns): records staged per path in LDS, one segment per path
  lp = engine.prepare(('heston', 0.03, 2.0, 0.04, 0.3, -0.7, 0.0), all_times,
                      keep, 366, np.array([np.log(100.0), 0.04]), n, rng,
                      torch.float32, mode='paths')
  lp.launch()
  t, _ = timed(lp.launch, 3)
  res['pathmajor_heston_ft'] = {
      'value': nbytes / t / 1e9, 'unit': 'GB/s', 'seconds': t,
      'bytes_written': nbytes, 'frac_hbm_peak': nbytes / t / HBM_PEAK,
      'path_steps_per_s': n * 365 / t,
      'note': 'contiguous [N][k][d] kernel mode; euler_sampling.sample '
              'returns the [N, k, d] view of the coalesced store '
              '(sample_call_heston_ft)'}
  del lp
  torch.cuda.empty_cache()
  # the drop-in call itself: euler_sampling.sample with HestonModel's
  # full-truncation coefficients (fused route), 2^20 paths, all 366 dates,
  # shifted Sobol draws; the whole call is timed (grid, tables, output
  # allocation, launch), and returns an [N, 366, 2] view of [366][2][N]
  hfn = tqf.models.HestonModel(2.0, 0.04, 0.3, -0.7).euler_ft_fns(0.03)
  sc = lambda: tqf.models.euler_sampling.sample(
      2, hfn[0], hfn[1], all_times[1:], num_time_steps=365, num_samples=n,
      initial_state=np.array([np.log(100.0), 0.04], np.float32),
      random_type=tqf.math.random.RandomType.SOBOL, seed=7, times_grid=None,
      dtype=torch.float32)
  sc()
  t, _ = timed(sc, 3)
  nb = 365 * 2 * n * 4
  res['sample_call_heston_ft'] = {
      'value': nb / t / 1e9, 'unit': 'GB/s (output of the whole call)',
      'seconds': t, 'bytes_written': nb, 'frac_hbm_peak': nb / t / HBM_PEAK,
      'path_steps_per_s': n * 365 / t}
  torch.cuda.empty_cache()
  # sobol.sample(730, 2^20) float32 (the C2 draw set, 3.06 GB)
  tsobol.sample(730, n)
  t, _ = timed(lambda: tsobol.sample(730, n), 3)
  nbytes = 730 * n * 4
  res['sobol_sample_730x2e20'] = {
      'value': nbytes / t / 1e9, 'unit': 'GB/s (output)', 'seconds': t,
      'frac_hbm_peak': nbytes / t / HBM_PEAK}
  torch.cuda.empty_cache()
  # HestonModel.sample_paths (QE), 2^20 paths, 365 steps to T = 1
  hm = tqf.models.HestonModel(2.0, 0.04, 0.3, -0.7)
  qe = lambda: hm.sample_paths([1.0], [np.log(100.0), 0.04], num_samples=n,
                               num_time_steps=365, seed=5)
  qe()
  t, _ = timed(qe, 3)
  res['heston_qe_sample_paths'] = {'value': n * 365 / t,
                                   'unit': 'path-steps/s', 'seconds': t}
  # GeometricBrownianMotion.sample_paths (exact), 2^20 paths x 100 dates
  gm = tqf.models.GeometricBrownianMotion(0.05, 0.2)
  gt = np.linspace(0.01, 1.0, 100)
  ge = lambda: gm.sample_paths(gt, initial_state=100.0, num_samples=n,
                               seed=42)
  ge()
  t, _ = timed(ge, 3)
  res['gbm_exact_sample_paths'] = {'value': n * 100 / t,
                                   'unit': 'path-steps/s', 'seconds': t,
                                   'output_GB_per_s': n * 100 * 4 / t / 1e9}
  torch.cuda.empty_cache()
  # Halton, 100 dimensions x 2^20 points (plain)
  ha = lambda: halton.sample(100, num_results=n, randomized=False)
  ha()
  t, _ = timed(ha, 3)
  res['halton_100x2e20'] = {'value': 100 * n / t, 'unit': 'point-dims/s',
                            'seconds': t}
  # pricing platform American options: 4 options, 2^20 paths, 100 dates
  spots = [90.0, 100.0, 110.0, 100.0]
  exp_t = [1.0, 1.0, 1.0, 2.0]

  def am():
    return american_option.bs_lsm_price(
        spots, exp_t, [100.0] * 4, [0.2, 0.25, 0.3, 0.2],
        [float(np.exp(-0.05 * e)) for e in exp_t], num_samples=n,
        num_exercise_times=100, is_call_option=False, seed=(1, 2))
  am()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  pr = am().cpu().numpy()
  t = time.perf_counter() - t0
  res['american_bs_lsm_4x2e20'] = {'value': 4 / t, 'unit': 'options/s',
                                   'seconds': t, 'prices': pr.tolist()}
  out['paths_and_points'] = res


FLOP_BASKET16 = 384  # FP32 per path-step of 16 assets, forward


def _basket16(out, timed, greeks, pool):
  # 16-asset basket with adjoint Greeks
  d = 16
  rs = np.random.RandomState(3)
  a = rs.standard_normal((d, d))
  c = a @ a.T + np.eye(d)
  corr = c / np.sqrt(np.outer(np.diag(c), np.diag(c)))
  sig = rs.uniform(0.15, 0.45, d)
  t, _ = timed(lambda: greeks.basket_call_greeks(
      np.full(d, 100.0), sig, corr, 0.02, 1.0, 100.0, 64, 262144), 3)
  ach = FLOP_BASKET16 * 262144 * 64 / t
  out['basket16_greeks'] = {
      'value': 262144 * 64 / t, 'unit': 'path-steps/s', 'seconds': t,
      'greeks': 33,
      'roofline': {'bound': 'valu', 'achieved': ach / 1e12,
                   'peak': PEAK_FP32 / 1e12, 'unit': 'TFLOP/s',
                   'frac': ach / PEAK_FP32, 'kernel': 'basket_kernel',
                   'flop_per_path_step': FLOP_BASKET16,
                   'note': 'forward FLOP only; the adjoint work of the 33 '
                           'Greeks is not counted'}}
  if pool is not None:
    out['basket16_greeks']['cpu_baseline'] = cpu_baseline_basket(
        pool, np.full(d, 100.0), sig, corr)


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=20)
  ap.add_argument('--warmup', type=int, default=3)
  ap.add_argument('--paths', type=int, default=2**20)
  ap.add_argument('--cpu-paths', type=int, default=2**20,
                  help='paths of the headline CPU baseline (default: the '
                       'full workload)')
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--no-check', action='store_true')
  ap.add_argument('--quick', action='store_true',
                  help='headline workload only (no secondary metrics)')
  ap.add_argument('--only', choices=SECONDARY, default=None,
                  help='run only this secondary workload and print its '
                       'result (for profiling)')
  ap.add_argument('--calib-copy', type=int, default=0, metavar='MIB',
                  help='only run 5 device-to-device copies of MIB MiB (the '
                       'known-size kernel that scales HBM counters)')
  args = ap.parse_args()
  if args.gpus != 1:
    print('bench.py measures one GPU: the engine runs on a single MI355X '
          'and has no multi-GPU path; run it with --gpus 1.',
          file=sys.stderr)
    sys.exit(2)

  # the CPU workers are spawned before this process touches the GPU
  pool = None if (args.no_cpu or args.calib_copy) else CpuPool(cpu_cores())
  try:
    run(args, pool)
  finally:
    if pool is not None:
      pool.__exit__(None, None, None)


def run(args, pool):
  import torch
  if args.calib_copy:
    n = args.calib_copy << 18  # float32 elements
    src = torch.ones(n, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(5):
      dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({'calib_copy_bytes_read': n * 4,
                      'calib_copy_bytes_written': n * 4}))
    return
  from tqf.engine import pricing

  dev = torch.device('cuda', 0)
  torch.cuda.set_device(dev)
  if args.only:
    print(json.dumps({args.only: secondary_metrics(args.only,
                                                   pool)[args.only]}))
    return
  launch = pricing.heston_ft_calls(num_paths=args.paths,
                                   num_time_steps=NUM_STEPS,
                                   maturities=MATURITIES, grid='reference',
                                   random='sobol_shifted', seed=7,
                                   dtype=torch.float32, prepare=True,
                                   **PARAMS)
  steps_per_path = int(launch.g.n_steps)
  # One step = path kernel (per-block payoff sums) + finalize, enqueued as
  # two calls so that events bracket the path kernel on its own stream.
  paths_call, finalize_call = launch.launch_split()
  for _ in range(args.warmup):
    paths_call()
    finalize_call()
  torch.cuda.synchronize()

  stream = torch.cuda.current_stream()
  ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
  ev_k = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  ev[0].record(stream)
  for i in range(args.steps):
    paths_call()
    ev_k[i].record(stream)
    finalize_call()
    ev[i + 1].record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  kern = [ev[i].elapsed_time(ev_k[i]) / 1e3 for i in range(args.steps)]
  avg = float(np.mean(kern))
  total_gpu = ev[0].elapsed_time(ev[-1]) / 1e3

  units = args.paths * steps_per_path
  value = units * args.steps / total_gpu
  flops = FLOP_PER_PATH_STEP * units
  achieved = flops / avg

  res = launch.result
  prices = res['prices'].cpu().numpy()
  ses = res['stderrs'].cpu().numpy()
  check = None
  if not args.no_check:
    check = accuracy_check(prices, ses)

  secondary = None
  if not args.quick:
    secondary = secondary_metrics(pool=pool)

  cpu = None
  if pool is not None:
    all_times, keep = pricing.heston_grid_for(MATURITIES, NUM_STEPS,
                                              'reference', torch.float32)
    shift = pricing.sobol_digital_shift(2 * (all_times.shape[0] - 1), 7)
    cpu = cpu_baseline_headline(pool, np.asarray(all_times), keep,
                                np.asarray(shift), args.cpu_paths)

  out = {
      'metric': METRIC,
      'value': value,
      'unit': 'path-steps/s',
      'n_gpus': 1,
      'steps': args.steps,
      'warmup': args.warmup,
      'ms_per_step': total_gpu / args.steps * 1e3,
      'higher_is_better': True,
      'scaling': 'weak',
      'vs_baseline': None,
      'dtype': 'f32',
      'data': 'synthetic',
      'config': {'workload': 'fused Heston full-truncation log-Euler, 2^20 '
                             'Sobol paths x 365 steps, 64 strikes x 4 '
                             'maturities, FP32 paths / FP64 accumulation',
                 'paths': args.paths, 'time_steps': steps_per_path,
                 'strikes': 64, 'maturities': list(MATURITIES),
                 'rng': 'sobol (Joe-Kuo, 32-digit shift, seed 7)'},
      'roofline': {'bound': 'valu', 'achieved': achieved / 1e12,
                   'peak': PEAK_FP32 / 1e12, 'unit': 'TFLOP/s',
                   'frac': achieved / PEAK_FP32,
                   'traffic': (TRAFFIC_BYTES_2POW20 if args.paths == 2**20
                               else None),
                   'kernel': 'euler_kernel<float, HESTON_FT, SOBOL_SHIFTED>',
                   'kernel_avg_ms': avg * 1e3,
                   'flop_per_path_step': FLOP_PER_PATH_STEP},
      'cpu_baseline': cpu,
      'check': check,
      'secondary': secondary,
      'wall_s': wall,
  }
  print(json.dumps(out))


if __name__ == '__main__':
  main()
