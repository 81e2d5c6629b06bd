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
FLOP_PER_PATH_STEP = 73  # algorithmic count, FP32
PEAK_FP32 = 157.3e12
PEAK_HBM = 8.0e12

PARAMS = dict(s0=100.0, v0=0.04, kappa=2.0, theta=0.04, xi=0.3, rho=-0.7,
              r=0.03)
MATURITIES = (0.25, 0.5, 0.75, 1.0)
NUM_STEPS = 365


def cpu_baseline(n_paths, chunk=2**14):
  """Times the oracle's numpy restatement (FP32 state) on host cores."""
  from oracle import heston as oheston
  from oracle import normal as onormal
  from tqf.engine import pricing
  import torch
  all_times, keep = pricing.heston_grid_for(MATURITIES, NUM_STEPS,
                                            'reference', torch.float32)
  steps = all_times.shape[0] - 1
  shift = pricing.sobol_digital_shift(2 * steps, 7)
  strikes = np.linspace(60, 140, 64)
  t0 = time.perf_counter()
  done = 0
  while done < n_paths:
    m = min(chunk, n_paths - done)
    ints = oheston.shifted_sobol_ints(2 * steps, m, shift, first=done)
    z = onormal.shifted_sobol_normals(ints).reshape(m, steps, 2)
    oheston.ft_euler_calls(np.log(PARAMS['s0']), PARAMS['v0'],
                           PARAMS['kappa'], PARAMS['theta'], PARAMS['xi'],
                           PARAMS['rho'], PARAMS['r'], strikes, all_times,
                           keep, z.astype(np.float32), dtype=np.float32)
    done += m
  dt = time.perf_counter() - t0
  return n_paths * steps / dt, dt


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=20)
  ap.add_argument('--warmup', type=int, default=3)
  ap.add_argument('--paths', type=int, default=2**20)
  ap.add_argument('--cpu-paths', type=int, default=2**17)
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--no-check', action='store_true')
  args = ap.parse_args()
  if args.gpus != 1:
    print('bench.py measures one GPU: this build runs on a single MI355X and '
          'has no multi-GPU path (see DESIGN.md, "More than one GPU").',
          file=sys.stderr)
    sys.exit(2)

  import torch
  from tqf.engine import pricing

  dev = torch.device('cuda', 0)
  torch.cuda.set_device(dev)
  launch = pricing.heston_ft_calls(num_paths=args.paths,
                                   num_time_steps=NUM_STEPS,
                                   maturities=MATURITIES, grid='reference',
                                   random='sobol_shifted', seed=7,
                                   dtype=torch.float32, prepare=True,
                                   **PARAMS)
  steps_per_path = int(launch.g.n_steps)
  for _ in range(args.warmup):
    launch.launch()
  torch.cuda.synchronize()

  stream = torch.cuda.current_stream()
  ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  ev[0].record(stream)
  for i in range(args.steps):
    launch.launch()
    ev[i + 1].record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  per = [ev[i].elapsed_time(ev[i + 1]) / 1e3 for i in range(args.steps)]
  avg = float(np.mean(per))
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
    from oracle import heston as oheston
    strikes = np.linspace(60, 140, 64)
    ref = np.array([[oheston.european_call(PARAMS['s0'], k, t, PARAMS['r'],
                                           PARAMS['v0'], PARAMS['kappa'],
                                           PARAMS['theta'], PARAMS['xi'],
                                           PARAMS['rho'])
                     for k in strikes] for t in MATURITIES])
    check = {
        'vs': 'semi-analytic Heston price (discretisation bias included)',
        'max_abs_err': float(np.max(np.abs(prices - ref))),
        'max_rel_err_price_gt_1': float(np.max(
            np.abs(prices - ref)[ref > 1] / ref[ref > 1])),
        'max_stderr': float(np.max(ses)),
    }

  cpu = None
  if not args.no_cpu:
    rate, secs = cpu_baseline(args.cpu_paths)
    cpu = {'value': rate, 'unit': 'path-steps/s', 'cores': 1, 'kind': 'port',
           'sample': '%d paths x %d steps, numpy FP32 restatement of the '
                     'same scheme incl. Sobol normals (%.1f s)' %
                     (args.cpu_paths, steps_per_path, secs)}

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
                   'frac': achieved / PEAK_FP32, 'traffic': None,
                   'kernel_avg_ms': avg * 1e3,
                   'flop_per_path_step': FLOP_PER_PATH_STEP},
      'cpu_baseline': cpu,
      'check': check,
      'wall_s': wall,
  }
  print(json.dumps(out))


if __name__ == '__main__':
  main()
