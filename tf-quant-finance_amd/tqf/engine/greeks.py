"""Adjoint Greeks behind torch.autograd.Function.

The reference computes Greeks by differentiating the sampler with
tf.GradientTape or tff.math.fwd_gradient (tff/math/gradient.py:22-235,
forward Jacobians through custom_loops.for_loop).  Here one kernel launch
returns the price and its pathwise adjoint derivatives; the autograd
Functions hand those to torch's backward, so `torch.autograd.grad(price,
params)` gives the Greeks without recording a graph of the simulation.
"""
import ctypes

import numpy as np
import torch

from tqf import _native
from tqf.engine import batch as batch_lib

__all__ = ['basket_call', 'BasketCall', 'HestonCall', 'GbmCall',
           'basket_call_greeks']


# per (device, stream, padded dimension, workgroups): the workspace, the
# output, and the parameter upload buffers (device and pinned host) with the
# event of the last upload.  Calls on one stream reuse them in stream order;
# the pinned buffer is rewritten only after its previous upload completed.
_BASKET_BUF = {}


def _basket_buffers(dev, stream, d4, n_blocks):
  key = (str(dev), stream, d4, n_blocks)
  b = _BASKET_BUF.get(key)
  if b is None:
    n_prm = 3 * d4 + d4 * d4
    b = {'ws': torch.empty(_native.lib().tqf_basket_workspace_size(
             d4, n_blocks) // 8 + 1, dtype=torch.float64, device=dev),
         'out': torch.empty(2 * (2 + 2 * d4), dtype=torch.float64,
                            device=dev),
         'prm': torch.empty(n_prm, dtype=torch.float32, device=dev),
         'prm_host': torch.empty(n_prm, dtype=torch.float32,
                                 pin_memory=True),
         'uploaded': None}
    _BASKET_BUF[key] = b
  return b


def _basket_launch(s0, sigma, weights, chol, rate, maturity, strike, n_steps,
                   n_paths, seed, stream_id, greeks, dev, n_blocks=None):
  """Launches the basket kernel; returns (out, idx, ns): `out` holds the
  means [ns] then the standard errors [ns] of the padded layout, `idx`
  selects price, rho, the d deltas and the d vegas from either half.  `out`
  is reused by the next call on the stream: copy what you keep."""
  d = len(s0)
  d4 = (d + 3) // 4 * 4
  pad = d4 - d

  def padv(v, fill):
    return np.concatenate([np.asarray(v, np.float64), np.full(pad, fill)])
  L = np.eye(d4)
  L[:d, :d] = np.asarray(chol, np.float64)
  prm = np.concatenate([padv(s0, 1.0), padv(sigma, 0.0), padv(weights, 0.0),
                        L.reshape(-1)]).astype(np.float32)
  if n_blocks is None:
    n_blocks = min(max((n_paths + 255) // 256, 1),
                   torch.cuda.get_device_properties(dev).multi_processor_count
                   * 8)
  stream = _native.stream_ptr(dev)
  b = _basket_buffers(dev, stream.value, d4, n_blocks)
  if b['uploaded'] is not None:
    b['uploaded'].synchronize()
  b['prm_host'].numpy()[:] = prm
  b['prm'].copy_(b['prm_host'], non_blocking=True)
  b['uploaded'] = torch.cuda.Event()
  b['uploaded'].record()
  _native.call('tqf_basket_call', ctypes.c_void_p(b['prm'].data_ptr()),
               ctypes.c_int32(d4), ctypes.c_int32(n_steps),
               ctypes.c_double(maturity), ctypes.c_double(rate),
               ctypes.c_double(strike), ctypes.c_int64(n_paths),
               ctypes.c_uint64(seed), ctypes.c_uint32(stream_id),
               ctypes.c_int32(int(greeks)), ctypes.c_int32(n_blocks),
               ctypes.c_void_p(b['ws'].data_ptr()),
               ctypes.c_void_p(b['out'].data_ptr()), stream)
  ns = 2 + 2 * d4
  idx = ([0, 1] + list(range(2, 2 + d)) + list(range(2 + d4, 2 + d4 + d)))
  return b['out'], idx, ns


def basket_call_greeks(s0, sigma, corr, rate, maturity, strike, n_steps,
                       n_paths, weights=None, seed=3, stream_id=0):
  """Price, rho, deltas[d], vegas[d] (means and standard errors)."""
  dev = _native.device()
  d = len(s0)
  w = np.full(d, 1.0 / d) if weights is None else np.asarray(weights)
  chol = np.linalg.cholesky(np.asarray(corr, np.float64))
  out, idx, ns = _basket_launch(s0, sigma, w, chol, rate, maturity, strike,
                                n_steps, n_paths, seed, stream_id, True, dev)
  o = out.cpu().numpy()  # one copy back, then the selection on the host
  m, s = o[:ns][idx], o[ns:][idx]
  return {'price': (m[0], s[0]), 'rho': (m[1], s[1]),
          'delta': (m[2:2 + d], s[2:2 + d]),
          'vega': (m[2 + d:], s[2 + d:])}


class BasketCall(torch.autograd.Function):
  """price = BasketCall.apply(s0[d], sigma[d], rate[], cfg) with adjoint
  backward (cfg: dict with corr, maturity, strike, n_steps, n_paths, seed)."""

  @staticmethod
  def forward(ctx, s0, sigma, rate, cfg):
    dev = _native.device()
    d = s0.shape[0]
    w = np.full(d, 1.0 / d) if cfg.get('weights') is None else cfg['weights']
    chol = np.linalg.cholesky(np.asarray(cfg['corr'], np.float64))
    out, idx, _ = _basket_launch(s0.detach().cpu().numpy(),
                                 sigma.detach().cpu().numpy(), w, chol,
                                 float(rate), cfg['maturity'], cfg['strike'],
                                 cfg['n_steps'], cfg['n_paths'],
                                 cfg.get('seed', 3), cfg.get('stream_id', 0),
                                 True, dev)
    mean = out[idx]  # a copy: `out` is reused by the next launch
    ctx.save_for_backward(mean[2:2 + d].to(s0.dtype),
                          mean[2 + d:].to(sigma.dtype))
    ctx.rho = mean[1]
    return mean[0].to(device=s0.device, dtype=s0.dtype)

  @staticmethod
  def backward(ctx, g):
    delta, vega = ctx.saved_tensors
    dev = g.device
    return (g * delta.to(dev), g * vega.to(dev),
            (g * ctx.rho.to(dev)).to(g.dtype), None)


def basket_call(s0, sigma, rate, **cfg):
  return BasketCall.apply(s0, sigma, rate, cfg)


def _contract_launch(contract, n_paths, seed, dev):
  plan = batch_lib.BatchPlan([dict(contract, n_paths=int(n_paths))],
                             tile=1024, n_blocks=None, seed=seed, dev=dev)
  out = plan.launch()
  return out[0, :8], out[0, 8:], plan


class HestonCall(torch.autograd.Function):
  """price = HestonCall.apply(s0, v0, rate, kappa, theta, xi, rho, cfg):
  full-truncation Heston call with checkpointed adjoint Greeks."""

  @staticmethod
  def forward(ctx, s0, v0, rate, kappa, theta, xi, rho, cfg):
    dev = _native.device()
    c = dict(kind=batch_lib.KIND_HESTON, n_steps=int(cfg['n_steps']),
             s0=float(s0), strike=float(cfg['strike']),
             maturity=float(cfg['maturity']), rate=float(rate), sigma=0.0,
             v0=float(v0), kappa=float(kappa), theta=float(theta),
             xi=float(xi), rho=float(rho), stream_id=cfg.get('stream_id', 0))
    mean, _, _ = _contract_launch(c, cfg['n_paths'], cfg.get('seed', 11), dev)
    ctx.g = mean[1:8].clone()
    s0t = torch.as_tensor(s0)
    return mean[0].to(device=s0t.device, dtype=s0t.dtype)

  @staticmethod
  def backward(ctx, g):
    d = ctx.g.to(device=g.device, dtype=g.dtype)
    return tuple(g * d[i] for i in range(7)) + (None,)


class GbmCall(torch.autograd.Function):
  """price = GbmCall.apply(s0, sigma, rate, cfg): log-Euler GBM call."""

  @staticmethod
  def forward(ctx, s0, sigma, rate, cfg):
    dev = _native.device()
    c = dict(kind=batch_lib.KIND_GBM, n_steps=int(cfg['n_steps']),
             s0=float(s0), strike=float(cfg['strike']),
             maturity=float(cfg['maturity']), rate=float(rate),
             sigma=float(sigma), v0=0.0, kappa=0.0, theta=0.0, xi=0.0, rho=0.0,
             stream_id=cfg.get('stream_id', 0))
    mean, _, _ = _contract_launch(c, cfg['n_paths'], cfg.get('seed', 11), dev)
    ctx.g = mean[1:4].clone()
    s0t = torch.as_tensor(s0)
    return mean[0].to(device=s0t.device, dtype=s0t.dtype)

  @staticmethod
  def backward(ctx, g):
    d = ctx.g.to(device=g.device, dtype=g.dtype)
    return g * d[0], g * d[1], g * d[2], None
