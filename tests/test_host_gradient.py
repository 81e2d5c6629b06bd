"""tqf.math gradient functions (host logic, CPU torch): the reference's
gradient_test.py cases (tff/math/gradient_test.py:22-118)."""
import numpy as np
import pytest
import torch

from tqf import math as tmath


def _t():
  return torch.arange(1, 3, dtype=torch.float32)


def _func(t):
  return torch.stack([t, t**2, t**3], 0)


def test_forward_gradient():
  g = tmath.fwd_gradient(_func, _t())
  assert g.shape == (3, 2)
  np.testing.assert_allclose(g.numpy(), [[1., 1.], [2., 4.], [3., 12.]])
  # the graph-mode form: a tensor computed from x
  t = _t().requires_grad_(True)
  np.testing.assert_allclose(tmath.fwd_gradient(_func(t), t).detach().numpy(),
                             [[1., 1.], [2., 4.], [3., 12.]])
  # a direction other than ones
  g = tmath.fwd_gradient(_func, _t(), input_gradients=torch.tensor([1., 0.]))
  np.testing.assert_allclose(g.numpy(), [[1., 0.], [2., 0.], [3., 0.]])


def test_forward_unconnected_gradient():
  zeros = torch.zeros(2)
  func = lambda t: torch.stack([zeros, zeros, zeros], 0)
  g = tmath.fwd_gradient(func, _t(), unconnected_gradients='zero')
  assert g.shape == (3, 2)
  np.testing.assert_allclose(g.numpy(), np.zeros((3, 2)))
  assert tmath.fwd_gradient(func, _t()) is None


def test_backward_gradient():
  g = tmath.gradients(_func, _t())
  assert g.shape == (2,)
  np.testing.assert_allclose(g.numpy(), [6., 17.])
  t = _t().requires_grad_(True)
  np.testing.assert_allclose(tmath.gradients(_func(t), t).numpy(), [6., 17.])
  ga, gb = tmath.gradients(lambda a, b: (a * b).sum(), [_t(), _t() + 1])
  np.testing.assert_allclose(ga.numpy(), [2., 3.])
  np.testing.assert_allclose(gb.numpy(), [1., 2.])
  # output_gradients weight the outputs
  w = torch.tensor([[1., 0.], [0., 0.], [0., 1.]])
  np.testing.assert_allclose(
      tmath.gradients(_func, _t(), output_gradients=w).numpy(), [1., 12.])


def test_backward_unconnected_gradient():
  zeros = torch.zeros(2)
  func = lambda t: torch.stack([zeros, zeros, zeros], 0)
  g = tmath.gradients(func, _t(), unconnected_gradients='zero')
  assert g.shape == (2,)
  np.testing.assert_allclose(g.numpy(), [0., 0.])
  assert tmath.gradients(func, _t()) is None
  with pytest.raises(ValueError):
    tmath.gradients(torch.zeros(3), _t())


def test_make_val_and_grad_fn():
  minimum = torch.tensor([1.0, 1.0], dtype=torch.float64)
  scales = torch.tensor([2.0, 3.0], dtype=torch.float64)

  @tmath.make_val_and_grad_fn
  def quadratic(x):
    return torch.sum(scales * (x - minimum)**2)

  val, grad = quadratic(torch.tensor([2.0, 2.0], dtype=torch.float64))
  assert abs(float(val) - 5.0) < 1e-12
  np.testing.assert_allclose(grad.numpy(), [4.0, 6.0])
  val, grads = tmath.value_and_gradient(lambda a, b: (a * b).sum(),
                                        [_t(), _t() + 1])
  assert float(val) == 8.0 and len(grads) == 2
