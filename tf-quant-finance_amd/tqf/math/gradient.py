"""Gradients (mirror of tff.math.fwd_gradient / gradients /
value_and_gradient / make_val_and_grad_fn, tff/math/gradient.py:22-235).

torch is always eager.  `func_or_y` is a callable, or (the reference's
graph-mode form, gradient.py:100-105, 159-162) a tensor already computed
from `x` / `xs` with gradients recorded; any other tensor raises
ValueError as the reference does in eager mode (gradient.py:110-112).

`fwd_gradient` is the reference's double-reverse construction
(gradient.py:100-119): g(w) = w^T J is a function of the placeholder w,
and its reverse derivative in the direction u is J u.  It needs the
backward of every operation on the path to be differentiable in its
incoming gradient.  The engine's autograd Functions are: the sampler's
backward (`euler_sampling.sample` with `watch_params` on a fused model)
is itself a Function whose derivative in the incoming gradient is the
forward-tangent kernel, so J u comes from one tangent pass of the paths,
with no per-step graph.

unconnected_gradients: 'none' (the default, tf.UnconnectedGradients.NONE)
returns None for inputs the output does not depend on; 'zero' returns
zeros.
"""
import functools

import torch

__all__ = ['fwd_gradient', 'gradients', 'value_and_gradient',
           'make_val_and_grad_fn']


def _prepare_func(func_or_y):
  if callable(func_or_y):
    return func_or_y
  if torch.is_tensor(func_or_y) and func_or_y.requires_grad:
    return lambda *args: func_or_y
  raise ValueError('`func_or_y` should be a callable in eager mode or when '
                   '`tf.GradientTape` is used.')


def _prepare_args(xs):
  if isinstance(xs, (list, tuple)):
    return list(xs), True
  return [xs], False


def _zeroes(unconnected_gradients):
  u = unconnected_gradients
  if u is None:
    return False
  name = getattr(u, 'name', u)
  name = str(name).lower()
  if name not in ('none', 'zero'):
    raise ValueError('unconnected_gradients must be "none" or "zero"')
  return name == 'zero'


def _watched(x):
  """x as a leaf that records gradients (a tensor that already requires
  grad is used as it is, so closures over it stay connected)."""
  if not torch.is_tensor(x):
    x = torch.as_tensor(x)
  if x.requires_grad:
    return x
  return x.detach().requires_grad_(True)


def fwd_gradient(func_or_y, x, input_gradients=None, use_gradient_tape=False,
                 unconnected_gradients=None, name=None):
  """Directional derivative J(F) u of F = func(x) (shape of F(x)).

  input_gradients: the direction u, shape of x (default ones)."""
  del use_gradient_tape, name
  f = _prepare_func(func_or_y)
  zero = _zeroes(unconnected_gradients)
  with torch.enable_grad():
    xw = _watched(x)
    y = f(xw)
    if not y.requires_grad:
      return torch.zeros_like(y) if zero else None
    w = torch.ones_like(y, requires_grad=True)
    g = torch.autograd.grad(y, xw, grad_outputs=w, create_graph=True,
                            allow_unused=True)[0]
    if g is None or not g.requires_grad:
      # F does not depend on x (or only through non-differentiable ops)
      return torch.zeros_like(y) if zero else None
    u = torch.ones_like(xw) if input_gradients is None else torch.as_tensor(
        input_gradients, dtype=xw.dtype, device=xw.device)
    out = torch.autograd.grad(g, w, grad_outputs=u, allow_unused=True)[0]
  if out is None:
    return torch.zeros_like(y) if zero else None
  return out


def gradients(func_or_y, xs, output_gradients=None, use_gradient_tape=False,
              unconnected_gradients=None, name=None):
  """w^T J: the gradient of func(*xs) with respect to each of xs
  (output_gradients w, default ones).  A list when xs is a list."""
  del use_gradient_tape, name
  f = _prepare_func(func_or_y)
  xs, is_list = _prepare_args(xs)
  zero = _zeroes(unconnected_gradients)
  with torch.enable_grad():
    xw = [_watched(x) for x in xs]
    y = f(*xw)
    grads = _grad(y, xw, output_gradients, zero)
  return grads if is_list else grads[0]


def _grad(y, xw, output_gradients, zero):
  ys = list(y) if isinstance(y, (list, tuple)) else [y]
  if output_gradients is None:
    ws = [torch.ones_like(t) for t in ys]
  else:
    ws = (list(output_gradients) if isinstance(output_gradients, (list, tuple))
          else [output_gradients])
    ws = [torch.as_tensor(w, dtype=t.dtype, device=t.device)
          for w, t in zip(ws, ys)]
  live = [t for t in ys if t.requires_grad]
  if not live:
    grads = [None] * len(xw)
  else:
    keep = [i for i, t in enumerate(ys) if t.requires_grad]
    grads = list(torch.autograd.grad([ys[i] for i in keep],
                                     xw, grad_outputs=[ws[i] for i in keep],
                                     allow_unused=True))
  if zero:
    grads = [torch.zeros_like(x) if g is None else g
             for g, x in zip(grads, xw)]
  return grads


def value_and_gradient(f, xs, output_gradients=None, use_gradient_tape=False,
                       unconnected_gradients=None, name=None):
  """(f(*xs), gradients of f with respect to xs) (gradient.py:186-233)."""
  del use_gradient_tape, name
  f = _prepare_func(f)
  xs, is_list = _prepare_args(xs)
  zero = _zeroes(unconnected_gradients)
  with torch.enable_grad():
    xw = [_watched(x) for x in xs]
    y = f(*xw)
    grads = _grad(y, xw, output_gradients, zero)
  y = y.detach() if torch.is_tensor(y) else y
  return y, (grads if is_list else grads[0])


def make_val_and_grad_fn(value_fn):
  """Decorator: x -> (value_fn(x), its gradient) (gradient.py:236-262)."""
  @functools.wraps(value_fn)
  def val_and_grad(x):
    return value_and_gradient(value_fn, x)
  return val_and_grad
