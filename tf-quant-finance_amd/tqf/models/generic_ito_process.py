"""Generic Ito process (mirror of tff.models.GenericItoProcess).

  GenericItoProcess   tff/models/generic_ito_process.py:27-317

    dX_i = a_i(t, X) dt + sum_j b_ij(t, X) dW_j

`sample_paths` is Euler-Maruyama through `tqf.models.euler_sampling.sample`
(generic_ito_process.py:300-317 delegates the same way): coefficient
callables that carry an engine spec (those of the tqf model classes) run in
the fused tqf_euler kernel, any other callables in the per-step loop on the
GPU, and `watch_params` makes the result differentiable with respect to
those tensors and the initial state.  The finite-difference solvers
(`fd_solver_backward` / `fd_solver_forward`, generic_ito_process.py:319-843)
belong to the PDE subsystem, which this engine does not provide.
"""
from tqf.models import euler_sampling

__all__ = ['GenericItoProcess']


class GenericItoProcess:
  """Ito process defined by a drift and a volatility callable."""

  def __init__(self, dim, drift_fn, volatility_fn, dtype=None, name=None):
    """dim: the process dimension n >= 1.  drift_fn(t, x) returns
    batch_shape + [num_samples, n]; volatility_fn(t, x) returns
    batch_shape + [num_samples, n, n] (generic_ito_process.py:30-133)."""
    if dim < 1:
      raise ValueError('Dimension must be 1 or greater.')
    if drift_fn is None or volatility_fn is None:
      raise ValueError('Both drift and volatility functions must be supplied.')
    self._dim = dim
    self._drift_fn = drift_fn
    self._volatility_fn = volatility_fn
    self._dtype = dtype
    self._name = name or 'generic_ito_process'

  def dim(self):
    """The dimension of the process."""
    return self._dim

  def dtype(self):
    """The data type of process realizations."""
    return self._dtype

  def name(self):
    """The name to give to ops created by this class."""
    return self._name

  def drift_fn(self):
    """The drift callable a(t, x)."""
    return self._drift_fn

  def volatility_fn(self):
    """The volatility callable b(t, x)."""
    return self._volatility_fn

  def sample_paths(self, times, num_samples=1, initial_state=None,
                   random_type=None, seed=None, swap_memory=True, name=None,
                   time_step=None, num_time_steps=None, skip=0,
                   precompute_normal_draws=True, times_grid=None,
                   normal_draws=None, watch_params=None,
                   validate_args=False):
    """Paths [num_samples, k, dim] at `times` by Euler sampling
    (generic_ito_process.py:187-317; arguments as euler_sampling.sample)."""
    del name
    return euler_sampling.sample(
        dim=self._dim, drift_fn=self._drift_fn,
        volatility_fn=self._volatility_fn, times=times,
        num_samples=num_samples, initial_state=initial_state,
        random_type=random_type, time_step=time_step,
        num_time_steps=num_time_steps, seed=seed, swap_memory=swap_memory,
        skip=skip, precompute_normal_draws=precompute_normal_draws,
        times_grid=times_grid, normal_draws=normal_draws,
        watch_params=watch_params, validate_args=validate_args,
        dtype=self._dtype)

  def fd_solver_backward(self, *args, **kwargs):
    """The PDE solvers (generic_ito_process.py:319-632) are not part of
    this engine."""
    raise NotImplementedError(
        'finite-difference solvers are not provided by the tqf engine')

  def fd_solver_forward(self, *args, **kwargs):
    """The PDE solvers (generic_ito_process.py:634-843) are not part of
    this engine."""
    raise NotImplementedError(
        'finite-difference solvers are not provided by the tqf engine')
