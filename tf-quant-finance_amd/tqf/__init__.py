"""tqf: MI355X-native Monte Carlo path simulation with tf-quant-finance's API.

Mirrors of the reference call surface:
  tqf.math.random.sobol.sample            <- tff.math.random.sobol.sample
  tqf.math.random.mv_normal_sample        <- tff.math.random.mv_normal_sample
  tqf.math.qmc.sobol_sample (+ nets)      <- tff.math.qmc
  tqf.models.euler_sampling.sample        <- tff.models.euler_sampling.sample
  tqf.models.GenericItoProcess            <- tff.models.GenericItoProcess
  tqf.math.fwd_gradient / gradients / value_and_gradient
                                          <- tff.math (gradient.py)
  tqf.models.GeometricBrownianMotion      <- tff.models.GeometricBrownianMotion
  tqf.models.MultivariateGeometricBrownianMotion
  tqf.models.HestonModel                  <- tff.models.HestonModel
  tqf.models.longstaff_schwartz           <- tff.models.longstaff_schwartz
  tqf.models.LocalVolatilityModel         <- tff.experimental.local_volatility
  tqf.experimental.american_option.bs_lsm_price
                                          <- pricing platform american_option
All compute runs in hand-written HIP kernels (libtqf.so) on the GPU.
"""
from tqf import errors
from tqf import math
from tqf import models
from tqf import engine
from tqf import experimental

__all__ = ['errors', 'math', 'models', 'engine', 'experimental']
