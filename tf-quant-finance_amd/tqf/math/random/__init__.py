"""Random and quasi-random normal draws (tff/math/random_ops/__init__.py)."""
from tqf.math.random import halton
from tqf.math.random import sobol
from tqf.math.random.multivariate_normal import multivariate_normal as mv_normal_sample
from tqf.math.random.multivariate_normal import RandomType

__all__ = ['halton', 'sobol', 'mv_normal_sample', 'RandomType']
