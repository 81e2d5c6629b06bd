"""Engine-level APIs beyond the reference's call surface (fused pricing)."""
from tqf.engine import euler
from tqf.engine import pricing

__all__ = ['euler', 'pricing']
