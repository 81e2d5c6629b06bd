"""Math foundations on the simulation path (random numbers, QMC,
piecewise-constant functions, gradients)."""
from tqf.math import piecewise
from tqf.math import qmc
from tqf.math import random
from tqf.math.gradient import fwd_gradient
from tqf.math.gradient import gradients
from tqf.math.gradient import make_val_and_grad_fn
from tqf.math.gradient import value_and_gradient

__all__ = ['piecewise', 'qmc', 'random', 'fwd_gradient', 'gradients',
           'make_val_and_grad_fn', 'value_and_gradient']
