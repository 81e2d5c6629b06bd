"""Path-simulation models (mirror of tff.models on the simulation path)."""
from tqf.models import euler_sampling
from tqf.models import longstaff_schwartz
from tqf.models import utils
from tqf.models.generic_ito_process import GenericItoProcess
from tqf.models.geometric_brownian_motion import GeometricBrownianMotion
from tqf.models.geometric_brownian_motion import MultivariateGeometricBrownianMotion
from tqf.models.heston import HestonModel
from tqf.models.local_volatility import LocalVolatilityModel
from tqf.models.local_volatility import LocalVolSurface

__all__ = ['euler_sampling', 'longstaff_schwartz', 'utils',
           'GenericItoProcess', 'GeometricBrownianMotion',
           'MultivariateGeometricBrownianMotion', 'HestonModel',
           'LocalVolatilityModel', 'LocalVolSurface']
