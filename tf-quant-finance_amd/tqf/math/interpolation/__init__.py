"""Interpolation used to build local volatility surfaces
(tff/math/interpolation): cubic splines, linear interpolation and the 2-D
spline-then-linear scheme.  Torch on the tensors' own device; these run
once when a surface is built, not per path."""
from tqf.math.interpolation import cubic
from tqf.math.interpolation import interpolation_2d
from tqf.math.interpolation import linear

__all__ = ['cubic', 'interpolation_2d', 'linear']
