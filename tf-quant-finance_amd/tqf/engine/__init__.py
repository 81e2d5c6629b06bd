"""Engine-level APIs beyond the reference's call surface: fused pricing,
batched heterogeneous pricing, adjoint Greeks."""
from tqf.engine import batch
from tqf.engine import euler
from tqf.engine import greeks
from tqf.engine import pricing

__all__ = ['batch', 'euler', 'greeks', 'pricing']
