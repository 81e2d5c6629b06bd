"""One-factor Hull-White short-rate model and the Bermudan swaption priced
by Longstaff-Schwartz on its simulated paths (tff.models.hull_white)."""
from tqf.models.hull_white.one_factor import HullWhiteModel1F
from tqf.models.hull_white.swaption import bermudan_swaption_price

__all__ = ['HullWhiteModel1F', 'bermudan_swaption_price']
