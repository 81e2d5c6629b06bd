"""Math foundations on the simulation path (random numbers, QMC)."""
from tqf.math import qmc
from tqf.math import random

__all__ = ['qmc', 'random']
