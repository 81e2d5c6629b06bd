"""Batched heterogeneous pricing (tqf_batch_price) with an autotuner and
hipGraph capture.

A batch is a list of European-call contracts, each GBM or Heston with its
own parameters, path count and step count.  `BatchPlan` splits the
contracts into work items of `tile` paths, orders the items longest first
for the persistent kernel, and keeps every device table resident, so a
launch is a single enqueue (and can be captured in a hipGraph).
`autotune` times candidate (tile, grid) shapes and caches the winner.
"""
import ctypes
import json
import os

import numpy as np
import torch

from tqf import _native

__all__ = ['BatchPlan', 'autotune', 'random_contracts', 'GREEK_NAMES',
           'KIND_GBM', 'KIND_HESTON', 'FLOP_PER_PATH_STEP',
           'FLOP_PER_PATH_STEP_TANGENT']

KIND_GBM, KIND_HESTON = 0, 1
GREEK_NAMES = {
    KIND_GBM: ['price', 'delta', 'vega', 'rho'],
    KIND_HESTON: ['price', 'delta', 'dv0', 'rho', 'dkappa', 'dtheta', 'dxi',
                  'drho'],
}
# FLOP per path-step of batch_kernel (csrc/batch.hip), counted with add/
# sub/mul/max/select 1, FMA 2, sqrt/log/exp/sin/cos/div 1, integer work 0,
# a Box-Muller normal 5 (10 per pair: 2 scale, log, mul, sqrt, sin, cos,
# 2 mul, amortised):
#   GBM     normal 5 + dw 1 + drift add 1 + sigma*dw FMA 2 + vega sum 1 = 10
#   Heston, reverse mode: forward 30 (2 normals 10; max, sqrt, 2 dw;
#           log-spot 6; variance 10) + the same 30 again when each
#           checkpoint segment is recomputed + reverse sweep 39 (max, sqrt,
#           2 dw 4; sqrt-bar 6; v-bar 7; rate-bar 2; kappa-bar 4; theta-bar
#           2; xi-bar 7; rho-bar 5; v update 2) = 99
#   Heston, tangent mode: forward 26 (2 normals 10; max, sqrt, 2 dw 4;
#           log-spot 6; variance 6 with the shared noise term) + shared
#           noise term 3 and 1/(2 sqrt v) 2 + five tangents x 10 (select,
#           scale, two FMA pairs) + parameter terms 14 (kappa 3, theta 1,
#           xi 5, rho 5) = 95
# Per-path payoff and Greek assembly (a few dozen FLOP per path) is left
# out: at >= 50 steps it is under 1% of the total.
FLOP_PER_PATH_STEP = {KIND_GBM: 10, KIND_HESTON: 99}
FLOP_PER_PATH_STEP_TANGENT = {KIND_GBM: 10, KIND_HESTON: 95}
GREEKS_ADJOINT, GREEKS_TANGENT = 0, 1
_FIELDS = ['kind', 'n_steps', 'n_paths', 's0', 'strike', 'maturity', 'rate',
           'sigma', 'v0', 'kappa', 'theta', 'xi', 'rho']
CACHE_PATH = os.environ.get(
    'TQF_AUTOTUNE_CACHE',
    os.path.join(os.path.expanduser('~'), '.cache', 'tqf', 'autotune.json'))


def random_contracts(n=4096, seed=20240607, rate=0.03):
  """Heterogeneous contract table: model Bernoulli(1/2) GBM/Heston,
  S0 ~ U[50,150], sigma or sqrt(v0) ~ U[0.1,0.5], kappa ~ U[0.5,5],
  theta ~ U[0.01,0.09], xi ~ U[0.1,0.8], rho ~ U[-0.9,0], K/S0 ~ U[0.7,1.3],
  T ~ U[0.1,3], paths = round(2^U[12,18]), steps ~ U{50..500}."""
  rs = np.random.RandomState(seed)
  kind = (rs.uniform(size=n) < 0.5).astype(np.int64)
  s0 = rs.uniform(50, 150, n)
  vol = rs.uniform(0.1, 0.5, n)
  kappa = rs.uniform(0.5, 5.0, n)
  theta = rs.uniform(0.01, 0.09, n)
  xi = rs.uniform(0.1, 0.8, n)
  rho = rs.uniform(-0.9, 0.0, n)
  moneyness = rs.uniform(0.7, 1.3, n)
  t = rs.uniform(0.1, 3.0, n)
  paths = np.rint(2.0**rs.uniform(12, 18, n)).astype(np.int64)
  steps = rs.randint(50, 501, n)
  return [dict(kind=int(kind[i]), n_steps=int(steps[i]),
               n_paths=int(paths[i]), s0=float(s0[i]),
               strike=float(moneyness[i] * s0[i]), maturity=float(t[i]),
               rate=rate, sigma=float(vol[i]), v0=float(vol[i]**2),
               kappa=float(kappa[i]), theta=float(theta[i]), xi=float(xi[i]),
               rho=float(rho[i])) for i in range(n)]


def _cost(c, n_paths, greeks=GREEKS_ADJOINT):
  # relative work of a path: Heston pays two normals per step and, in
  # reverse mode, a recompute and a reverse sweep
  if c['kind'] != KIND_HESTON:
    return n_paths * c['n_steps']
  return n_paths * c['n_steps'] * (3.2 if greeks == GREEKS_ADJOINT else 2.2)


class BatchPlan:
  """Device-resident tables for one batch of contracts."""

  def __init__(self, contracts, tile=1024, n_blocks=None, use_counter=True,
               seed=11, dev=None, greeks='adjoint'):
    """greeks: 'adjoint' (reverse sweep over sqrt(steps) checkpoints) or
    'tangent' (forward tangents) for the Heston contracts; the same
    pathwise estimator either way."""
    self.dev = dev or _native.device()
    self.contracts = list(contracts)
    if greeks not in ('adjoint', 'tangent'):
      raise ValueError("greeks must be 'adjoint' or 'tangent'")
    self.greeks = greeks
    mode = GREEKS_TANGENT if greeks == 'tangent' else GREEKS_ADJOINT
    self.tile = int(tile)
    self.seed = int(seed)
    self.use_counter = bool(use_counter)
    if n_blocks is None:
      props = torch.cuda.get_device_properties(self.dev)
      n_blocks = props.multi_processor_count * 8
    self.n_blocks = int(n_blocks)
    n = len(self.contracts)
    table = (_native.Contract * n)()
    for i, c in enumerate(self.contracts):
      for f in _FIELDS:
        setattr(table[i], f, c[f])
      table[i].stream_id = int(c.get('stream_id', i))
      table[i].greeks_mode = mode
    items, first, count, costs = [], [], [], []
    for i, c in enumerate(self.contracts):
      first.append(len(items))
      k = 0
      for p0 in range(0, c['n_paths'], self.tile):
        m = min(self.tile, c['n_paths'] - p0)
        items.append((i, p0, m))
        costs.append(_cost(c, m, mode))
        k += 1
      count.append(k)
    it = (_native.WorkItem * len(items))()
    for j, (ci, p0, m) in enumerate(items):
      it[j].contract, it[j].first_path, it[j].n_paths = ci, p0, m
    order = np.argsort(-np.asarray(costs), kind='stable').astype(np.int32)
    self.n_items = len(items)
    self.max_steps = max(c['n_steps'] for c in self.contracts)

    def up(buf, nbytes):
      raw = np.frombuffer(buf, dtype=np.uint8, count=nbytes).copy()
      return torch.from_numpy(raw).to(self.dev)
    self._contracts = up(table, ctypes.sizeof(table))
    self._items = up(it, ctypes.sizeof(it))
    self._order = torch.from_numpy(order).to(self.dev)
    self._first = torch.as_tensor(np.asarray(first, np.int32), device=self.dev)
    self._count = torch.as_tensor(np.asarray(count, np.int32), device=self.dev)
    ws = _native.lib().tqf_batch_workspace_size(self.n_items, self.n_blocks,
                                                self.max_steps)
    self._ws = torch.empty(ws // 8 + 1, dtype=torch.float64, device=self.dev)
    self.out = torch.empty((n, 16), dtype=torch.float64, device=self.dev)
    # a tangent plan launches the kernel without the reverse sweep
    self.flags = ((_native.BATCH_COUNTER if self.use_counter else 0) |
                  (_native.BATCH_TANGENT_ONLY if mode == GREEKS_TANGENT
                   else 0))
    self.total_path_steps = int(sum(c['n_paths'] * c['n_steps']
                                    for c in self.contracts))
    flop = (FLOP_PER_PATH_STEP_TANGENT if mode == GREEKS_TANGENT else
            FLOP_PER_PATH_STEP)
    self.total_flop = int(sum(c['n_paths'] * c['n_steps'] * flop[c['kind']]
                              for c in self.contracts))

  def launch(self):
    _native.call('tqf_batch_price', ctypes.c_void_p(self._contracts.data_ptr()),
                 ctypes.c_int32(len(self.contracts)),
                 ctypes.c_void_p(self._items.data_ptr()),
                 ctypes.c_void_p(self._order.data_ptr()),
                 ctypes.c_int32(self.n_items),
                 ctypes.c_void_p(self._first.data_ptr()),
                 ctypes.c_void_p(self._count.data_ptr()),
                 ctypes.c_uint64(self.seed), ctypes.c_int32(self.n_blocks),
                 ctypes.c_int32(self.max_steps),
                 ctypes.c_int32(self.flags),
                 ctypes.c_void_p(self._ws.data_ptr()),
                 ctypes.c_void_p(self.out.data_ptr()),
                 _native.stream_ptr(self.dev))
    return self.out

  def capture(self):
    """Captures one launch in a hipGraph; returns the graph."""
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream(self.dev)
    s.wait_stream(torch.cuda.current_stream(self.dev))
    with torch.cuda.stream(s):
      self.launch()  # warm-up outside capture
    torch.cuda.current_stream(self.dev).wait_stream(s)
    torch.cuda.synchronize(self.dev)
    with torch.cuda.graph(g):
      self.launch()
    return g

  def results(self):
    """List of dicts of means and standard errors per contract."""
    o = self.out.cpu().numpy()
    res = []
    for i, c in enumerate(self.contracts):
      names = GREEK_NAMES[c['kind']]
      res.append({n: (o[i, q], o[i, 8 + q]) for q, n in enumerate(names)})
    return res


def _time(plan, reps):
  plan.launch()
  torch.cuda.synchronize(plan.dev)
  a = torch.cuda.Event(enable_timing=True)
  b = torch.cuda.Event(enable_timing=True)
  a.record()
  for _ in range(reps):
    plan.launch()
  b.record()
  torch.cuda.synchronize(plan.dev)
  return a.elapsed_time(b) / 1e3 / reps


def autotune(contracts, tiles=(512, 1024, 2048, 4096), blocks_per_cu=(4, 8),
             greeks=('adjoint', 'tangent'), reps=2, seed=11,
             cache_path=CACHE_PATH, key=None):
  """Times (tile, grid, Greek mode) candidates on this batch and caches the
  best.

  Returns (best_config dict, timings list).  The cache is keyed by the
  device name and `key` (default: a signature of the contract table).
  """
  dev = _native.device()
  name = torch.cuda.get_device_properties(dev).name
  cus = torch.cuda.get_device_properties(dev).multi_processor_count
  if key is None:
    key = 'n%d_ps%d' % (len(contracts), int(sum(c['n_paths'] * c['n_steps']
                                                  for c in contracts)))
  full_key = '%s|%s' % (name, key)
  cache = {}
  if cache_path and os.path.exists(cache_path):
    try:
      with open(cache_path) as f:
        cache = json.load(f)
    except (OSError, ValueError):
      cache = {}
  if full_key in cache:
    return cache[full_key], None
  timings = []
  for mode in greeks:
    for tile in tiles:
      for bpc in blocks_per_cu:
        plan = BatchPlan(contracts, tile=tile, n_blocks=cus * bpc, seed=seed,
                         dev=dev, greeks=mode)
        t = _time(plan, reps)
        timings.append({'tile': tile, 'n_blocks': cus * bpc, 'greeks': mode,
                        'seconds': t})
  best = min(timings, key=lambda r: r['seconds'])
  cfg = {'tile': best['tile'], 'n_blocks': best['n_blocks'],
         'greeks': best['greeks']}
  if cache_path:
    try:
      os.makedirs(os.path.dirname(cache_path), exist_ok=True)
      cache[full_key] = cfg
      with open(cache_path, 'w') as f:
        json.dump(cache, f, indent=1)
    except OSError:
      pass
  return cfg, timings
