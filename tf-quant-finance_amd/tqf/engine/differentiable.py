"""Differentiable fused paths: euler_sampling.sample with gradients.

The reference makes its sampler differentiable with respect to
`watch_params` and the initial state by carrying forward Jacobians through
the step loop (tff/models/euler_sampling.py:467-510 with
tff/math/custom_loops.py:20-214), so memory does not grow with the number
of steps.  Here the forward pass is the fused tqf_euler kernel and the
backward pass is the tqf_euler_tangent kernel, which re-simulates every
path (same normals, same update) with the tangents of its state in
registers:

  _EulerPaths      forward: paths [N, k, d];  backward: _EulerVjp
  _EulerVjp        forward: the vector-Jacobian product (gradients of the
                   inputs);  backward, in the incoming gradient: the
                   Jacobian-vector product (tangent paths), which is what
                   tqf.math.fwd_gradient's double-reverse construction asks
                   for (tff/math/gradient.py:100-119).

Inputs: the initial state (shared [d] or per path [N, d]) and the model's
parameters -- ('gbm', mu, sigma), ('const', a, b), ('heston', r, kappa,
theta, xi, rho, variant).  Second derivatives with respect to the inputs
are not provided.
"""
import ctypes

import torch

from tqf import _native
from tqf.engine import euler as engine

__all__ = ['TANGENT_KINDS', 'needs_gradients', 'sample_paths']

TANGENT_KINDS = {'gbm': 2, 'const': 2, 'heston': 5}
_TANGENT_RNG = ('draws', 'philox', 'philox_antithetic')


def _params(spec):
  return list(spec[1:1 + TANGENT_KINDS[spec[0]]])


def needs_gradients(spec, x0):
  """True when grad mode is on and the initial state or a parameter of the
  spec is a tensor that requires grad."""
  if not torch.is_grad_enabled() or spec[0] not in TANGENT_KINDS:
    return False
  vals = [x0] + _params(spec)
  return any(torch.is_tensor(v) and v.requires_grad for v in vals)


class _Info:
  """Everything but the differentiable inputs of one fused sampling call."""

  def __init__(self, spec, all_times, keep_mask, k, num_samples, rng, dtype,
               dev):
    self.spec, self.all_times, self.keep_mask = spec, all_times, keep_mask
    self.k, self.num_samples, self.rng = k, num_samples, rng
    self.dtype, self.dev = dtype, dev

  def launch(self, x0, params):
    vals = [float(p) for p in params]
    spec = (self.spec[0],) + tuple(vals) + tuple(
        self.spec[1 + len(vals):])
    return engine.prepare(spec, self.all_times, self.keep_mask, self.k,
                          x0.detach(), self.num_samples, self.rng, self.dtype,
                          mode='paths', dev=self.dev)


def _tangent(launch, mode, n_par, grad_paths=None, direction=None,
             dir_x0=None):
  d = int(launch.m.dim)
  n = launch.num_samples
  dev = launch.dev
  t = _native.Tangent()
  t.mode = mode
  keep = []
  res = {}
  if mode == _native.TANGENT_VJP:
    gp = torch.zeros(d + n_par, dtype=torch.float64, device=dev)
    ws = torch.empty(_native.lib().tqf_euler_tangent_workspace_size(n) // 8
                     + 1, dtype=torch.float64, device=dev)
    g = grad_paths.to(launch.dtype).contiguous()
    keep += [gp, ws, g]
    t.grad_paths, t.grad_params, t.workspace = (g.data_ptr(), gp.data_ptr(),
                                                ws.data_ptr())
    res['params'] = gp
    if launch.x0_stride:
      gx = torch.empty((n, d), dtype=launch.dtype, device=dev)
      t.grad_x0 = gx.data_ptr()
      res['x0'] = gx
  else:
    out = torch.empty((n, int(launch.g.n_slots), d),
                      dtype=launch.dtype, device=dev)
    dirt = direction.to(torch.float64).contiguous()
    keep.append(dirt)
    t.direction, t.tangent_paths = dirt.data_ptr(), out.data_ptr()
    if dir_x0 is not None:
      dx = dir_x0.to(launch.dtype).reshape(n, d).contiguous()
      keep.append(dx)
      t.dir_x0 = dx.data_ptr()
    res['paths'] = out
  _native.call('tqf_euler_tangent', ctypes.byref(launch.m),
               ctypes.byref(launch.g), ctypes.byref(launch.r),
               ctypes.byref(t), ctypes.c_int64(n),
               ctypes.c_void_p(launch.x0_t.data_ptr()),
               ctypes.c_int64(launch.x0_stride),
               ctypes.c_int32(_native.dtype_code(launch.dtype)),
               _native.stream_ptr(dev))
  res['_keep'] = keep
  return res


def _x0_grad(x0, shared, per_path):
  """The gradient of the kernel's initial state mapped back to the shape
  the caller passed (a broadcast input receives the sum)."""
  if per_path is not None:
    return per_path.reshape(x0.shape).to(x0.dtype)
  d = shared.shape[0]
  if x0.numel() == d:
    return shared.reshape(x0.shape).to(x0.dtype)
  return shared.sum().reshape(x0.shape).to(x0.dtype)


class _EulerVjp(torch.autograd.Function):

  @staticmethod
  def forward(ctx, launch, n_par, grad_paths, x0, *params):
    r = _tangent(launch, _native.TANGENT_VJP, n_par, grad_paths=grad_paths)
    d = int(launch.m.dim)
    gp = r['params']
    ctx.launch, ctx.n_par = launch, n_par
    ctx.x0_shape, ctx.x0_numel = x0.shape, x0.numel()
    ctx.per_path = 'x0' in r
    gx0 = _x0_grad(x0, gp[:d], r.get('x0'))
    gps = [gp[d + i].to(p.dtype).reshape(p.shape) for i, p in
           enumerate(params)]
    return (gx0,) + tuple(gps)

  @staticmethod
  def backward(ctx, ggx0, *ggparams):
    # derivative of the VJP in the incoming gradient: the tangent paths in
    # the direction (ggx0, ggparams)
    launch, n_par = ctx.launch, ctx.n_par
    d = int(launch.m.dim)
    direction = torch.zeros(d + n_par, dtype=torch.float64,
                            device=launch.dev)
    dir_x0 = None
    if ggx0 is not None:
      if ctx.per_path:
        dir_x0 = ggx0
      elif ctx.x0_numel == d:
        direction[:d] = ggx0.reshape(d).to(torch.float64)
      else:
        direction[:d] = ggx0.reshape(()).to(torch.float64)
    for i, gg in enumerate(ggparams):
      if gg is not None:
        direction[d + i] = gg.reshape(()).to(torch.float64)
    r = _tangent(launch, _native.TANGENT_JVP, n_par, direction=direction,
                 dir_x0=dir_x0)
    return (None, None, r['paths'], None) + (None,) * len(ggparams)


class _EulerPaths(torch.autograd.Function):

  @staticmethod
  def forward(ctx, info, x0, *params):
    launch = info.launch(x0, params)
    out = launch.launch()
    ctx.launch = launch
    ctx.save_for_backward(x0, *params)
    return out

  @staticmethod
  def backward(ctx, g):
    x0, *params = ctx.saved_tensors
    grads = _EulerVjp.apply(ctx.launch, len(params), g, x0, *params)
    return (None,) + tuple(grads)


def sample_paths(spec, all_times, keep_mask, k, x0, num_samples, rng, dtype,
                 dev):
  """Paths [N, k, d] of a fused model, differentiable with respect to the
  initial state and the spec's parameters (tensors or numbers)."""
  if spec[0] not in TANGENT_KINDS:
    raise NotImplementedError('gradients of %s paths' % spec[0])
  if rng.kind not in _TANGENT_RNG:
    raise ValueError('the differentiable sampler takes draws or Philox')
  d = engine.model_dim(spec)
  x0_t = torch.as_tensor(x0, device=dev)
  if not x0_t.is_floating_point() or x0_t.dtype != dtype:
    x0_t = x0_t.to(dtype)
  if x0_t.dim() == 0:
    x0_t = x0_t.reshape(1)
  params = [p if torch.is_tensor(p) else torch.tensor(float(p), dtype=dtype,
                                                      device=dev)
            for p in _params(spec)]
  info = _Info(spec, all_times, keep_mask, k, int(num_samples), rng, dtype,
               dev)
  return _EulerPaths.apply(info, x0_t, *params)
