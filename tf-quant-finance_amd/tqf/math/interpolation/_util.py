"""Shared helpers: tensor conversion and batch broadcasting."""
import numpy as np
import torch


def as_tensor(x, dtype=None, like=None):
  """x as a torch tensor; dtype from `dtype`, else `like`, else float64 for
  Python / numpy floats."""
  if isinstance(dtype, np.dtype) or (isinstance(dtype, type) and
                                     issubclass(dtype, np.generic)):
    dtype = torch.float32 if np.dtype(dtype) == np.float32 else torch.float64
  dev = like.device if isinstance(like, torch.Tensor) else None
  if isinstance(x, torch.Tensor):
    x = x.to(device=dev or x.device)
    return x.to(dtype) if dtype is not None else x
  if dtype is None:
    dtype = like.dtype if isinstance(like, torch.Tensor) else torch.float64
  return torch.as_tensor(np.array(x, np.float64), dtype=dtype, device=dev)


def broadcast_batch(*tensors):
  """Broadcasts all but the last dimension of each tensor to one shape."""
  batch = torch.broadcast_shapes(*[t.shape[:-1] for t in tensors])
  return [t.expand(*batch, t.shape[-1]) for t in tensors]


def divide_no_nan(a, b):
  """a / b, 0 where b == 0 (with a finite gradient there)."""
  zero = b == 0
  return torch.where(zero, torch.zeros_like(a * b),
                     a / torch.where(zero, torch.ones_like(b), b))
