"""The allocation-free payoff entry point (tqf_euler_payoff_noalloc) against
tqf_euler with a caller workspace: two calls in a row, the second with
fewer paths, for the headline scheme (float32, shifted Sobol) and a
float64 Philox model; and the limits it enforces."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
  from tqf import _native
  from tqf.engine import euler as engine
  from tqf.engine import pricing


def _noalloc(launch, workspace=None, n_paths=None):
  o = _native.Output.from_buffer_copy(launch.o)
  o.workspace = workspace
  res = launch.result
  out = {k: torch.full_like(res[k], np.nan) for k in ('prices', 'stderrs',
                                                      'sums')}
  o.prices, o.stderrs, o.sums = (out['prices'].data_ptr(),
                                 out['stderrs'].data_ptr(),
                                 out['sums'].data_ptr())
  n = launch.num_samples if n_paths is None else n_paths
  _native.call('tqf_euler_payoff_noalloc', ctypes.byref(launch.m),
               ctypes.byref(launch.g), ctypes.byref(launch.r),
               ctypes.byref(o), ctypes.c_int64(n),
               ctypes.c_void_p(launch.x0_t.data_ptr()),
               ctypes.c_int64(launch.x0_stride),
               ctypes.c_int32(_native.dtype_code(launch.dtype)),
               _native.stream_ptr(launch.dev))
  return out


def _check(launch):
  want = {k: v.clone() for k, v in launch.launch().items()
          if k in ('prices', 'stderrs', 'sums')}
  got = _noalloc(launch)
  torch.cuda.synchronize()
  for k in want:
    assert torch.equal(got[k], want[k]), k


def test_noalloc_payoffs_equal_the_workspace_entry_point_twice():
  # the headline scheme (64 strikes x 4 maturities), then a smaller ragged
  # call: the second call's result must not see the first one's partials
  print('noalloc: headline scheme', flush=True)
  for n in (2**18, 2**13 + 77):
    _check(pricing.heston_ft_calls(num_paths=n, num_time_steps=73,
                                   prepare=True))
  print('noalloc: float64 philox GBM', flush=True)
  all_times, keep = pricing.heston_grid_for([0.5, 1.0], 50, 'reference',
                                            torch.float64)
  for n in (100000, 999):
    _check(pricing.european_calls(('gbm', 0.05, 0.2), [100.0],
                                  np.linspace(80, 120, 8), all_times, keep,
                                  n, engine.RngSpec('philox', seed=42),
                                  discount_rate=0.05, dtype=torch.float64,
                                  prepare=True))


def test_noalloc_rejects_a_workspace_and_sizes_beyond_its_scratch():
  launch = pricing.heston_ft_calls(num_paths=2**12, num_time_steps=8,
                                   prepare=True)
  with pytest.raises(ValueError, match='workspace must be NULL'):
    _noalloc(launch, workspace=launch.o.workspace)
  # 2^24 paths x 256 payoffs need 68 MB of partials (no launch happens)
  with pytest.raises(ValueError, match='need'):
    _noalloc(launch, n_paths=2**24)
