"""Callers of the simulation path from tff.experimental (American options
priced by Longstaff-Schwartz on shared GBM draws)."""
from tqf.experimental import american_option

__all__ = ['american_option']
